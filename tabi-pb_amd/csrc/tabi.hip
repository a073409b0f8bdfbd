// MI355X (gfx950) treecode-accelerated boundary-integral Poisson-Boltzmann solver.
//
// Device data live in tree order (faces sorted into leaf order by the tree
// build), so every target group of 64 consecutive faces -- one wavefront -- is a
// spatially compact patch, and every leaf is a contiguous source range.
//
// Per matvec (y = A x, TREE_COMPP_PB, src/treecode3d_pb.f:1130-1189):
//   k_mom_raw / k_mom_final  cluster moments of the 4 distinct source-charge sets
//                            (area*dphi/dn and n_d*area*phi), folded with the
//                            derivative factors into 8 contraction vectors
//   k_far<P>                 particle-cluster Taylor far field, one wave per task
//   k_near                   direct near field, one wave per task, sources read
//                            wave-uniformly through the scalar cache
//   k_reduce                 y = pre*x - (ordered sum of the task partials)
// Per preconditioner solve (LEAFMATVECpara, src/treecode3d_pb.f:1393-1542):
//   k_lu_solve               blocked forward/back substitution with the leaf
//                            factors computed once at set-up
// Set-up: k_lists (per-target MAC traversal with lane masks), k_lu_assemble,
//   k_lu_panel / k_lu_linv / k_lu_swap_trsm / k_lu_strip_trsm / k_lu_gemm
//   (batched two-level blocked LU with partial pivoting over all leaf blocks
//   at once).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <array>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tabi_hip.h"
#include "tabi_internal.h"
#include "taylor.h"

using tabi::HostTree;
using u64 = unsigned long long;

namespace {

constexpr double kPi = 3.141592653589793238462643383279;
constexpr double kOneOver4Pi = 0.25 / kPi;
constexpr int kWave = 64;
constexpr int kMaxOrder = 6;
constexpr int kStack = 512;      // per-wave traversal stack (entries)
constexpr int kLuNB = 32;        // LU panel width

thread_local std::string g_err;
thread_local double g_times[8] = {0, 0, 0, 0, 0, 0, 0, 0};
thread_local double g_many[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw Error(std::string(#expr) + " failed: " + hipGetErrorString(e_));     \
    } while (0)

// The stream whose work uses the device blocks this thread allocates and frees
// (the context's stream while the library works for a context; null outside,
// where the caller guarantees that no queued work uses a freed block).
thread_local hipStream_t g_pool_stream = nullptr;

struct PoolScope {
    hipStream_t prev;
    explicit PoolScope(hipStream_t s) : prev(g_pool_stream) { g_pool_stream = s; }
    ~PoolScope() { g_pool_stream = prev; }
    PoolScope(const PoolScope &) = delete;
    PoolScope &operator=(const PoolScope &) = delete;
};

// Caching device allocator: freed blocks are kept per size class and handed
// out again, so repeated solves of similar surfaces do not pay hipMalloc /
// hipFree (which map and unmap pages) each time.  Frees are stream-ordered:
// a block remembers the stream it was freed on (kernels queued there may
// still read it) and is handed out again at once only for work on that same
// stream, which runs after them; for work on another stream only once the
// freeing stream has drained (hipStreamQuery).  Contexts used from different
// host threads therefore never receive a block another thread's pending work
// still reads.  Memory is returned to the runtime by tabi_release_memory().
class DevicePool {
public:
    static DevicePool &get()
    {
        // never destroyed: contexts held in static storage return their blocks
        // during exit-time destruction, after function-local statics are gone
        static DevicePool *p = new DevicePool;
        return *p;
    }
    void *alloc(size_t bytes)
    {
        const size_t cls = size_class(bytes);
        const hipStream_t s = g_pool_stream;
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = free_.find(cls);
            if (it != free_.end()) {
                auto &v = it->second;
                int looked = 0;
                for (size_t k = v.size(); k-- > 0;) {
                    const hipStream_t fs = v[k].s;
                    bool ok = fs == nullptr || fs == s;
                    if (!ok && looked < 8) {
                        ++looked;
                        ok = hipStreamQuery(fs) == hipSuccess;
                        if (ok) v[k].s = nullptr;   // drained: safe for any stream
                    }
                    if (ok) {
                        void *p = v[k].p;
                        v.erase(v.begin() + (std::ptrdiff_t)k);
                        return p;
                    }
                }
            }
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, cls);
        if (e != hipSuccess) {
            release_all();
            HIPCHK(hipMalloc(&p, cls));
        }
        std::lock_guard<std::mutex> g(mu_);
        size_of_[p] = cls;
        return p;
    }
    void free(void *p)
    {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        free_[size_of_[p]].push_back({p, g_pool_stream});
    }
    // The stream is about to be destroyed after a synchronisation: its blocks
    // have no pending users.
    void retire(hipStream_t s)
    {
        if (!s) return;
        std::lock_guard<std::mutex> g(mu_);
        for (auto &kv : free_)
            for (auto &b : kv.second)
                if (b.s == s) b.s = nullptr;
    }
    void release_all()
    {
        std::lock_guard<std::mutex> g(mu_);
        (void)hipDeviceSynchronize();
        for (auto &kv : free_)
            for (auto &b : kv.second) {
                (void)hipFree(b.p);
                size_of_.erase(b.p);
            }
        free_.clear();
    }

private:
    struct Block {
        void *p;
        hipStream_t s;   // stream the block was freed on (null: no pending user)
    };
    static size_t size_class(size_t b)
    {
        if (b < 256) return 256;
        // 1/8-octave classes: at most 12.5% rounding waste
        size_t top = 1;
        while ((top << 1) <= b) top <<= 1;
        const size_t step = std::max<size_t>(top / 8, 256);
        return (b + step - 1) / step * step;
    }
    std::mutex mu_;
    std::map<size_t, std::vector<Block>> free_;
    std::map<void *, size_t> size_of_;
};

// Synchronises a stream, marks its cached blocks free of pending users and
// destroys it.
void destroy_stream(hipStream_t s)
{
    if (!s) return;
    (void)hipStreamSynchronize(s);
    DevicePool::get().retire(s);
    (void)hipStreamDestroy(s);
}

// Streams (default priority) for one-shot solves, kept per device and handed
// out again: creating and destroying a solve's two streams each time cost
// ~0.3 ms per solve.  A stream is synchronised and its cached blocks marked
// free of pending users when it comes back.
class StreamPool {
public:
    static hipStream_t acquire()
    {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> g(mu());
            auto &v = free_list()[dev];
            if (!v.empty()) {
                hipStream_t s = v.back();
                v.pop_back();
                return s;
            }
        }
        hipStream_t s = nullptr;
        HIPCHK(hipStreamCreate(&s));
        return s;
    }
    static void release(hipStream_t s)
    {
        if (!s) return;
        (void)hipStreamSynchronize(s);
        DevicePool::get().retire(s);
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) {
            (void)hipStreamDestroy(s);
            return;
        }
        std::lock_guard<std::mutex> g(mu());
        free_list()[dev].push_back(s);
    }

private:
    // never destroyed (streams held in static storage outlive the runtime's
    // own teardown otherwise)
    static std::mutex &mu() { static std::mutex *m = new std::mutex; return *m; }
    static std::map<int, std::vector<hipStream_t>> &free_list()
    {
        static auto *f = new std::map<int, std::vector<hipStream_t>>;
        return *f;
    }
};

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        if (count == 0) count = 1;
        p = static_cast<T *>(DevicePool::get().alloc(count * sizeof(T)));
        n = count;
    }
    void release()
    {
        if (p) DevicePool::get().free(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s)
    {
        alloc(count);
        HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

}  // namespace

namespace tabi {
void *device_alloc(size_t bytes) { return DevicePool::get().alloc(bytes); }
void device_free(void *p) { DevicePool::get().free(p); }
}  // namespace tabi

namespace {

// An input array as a device pointer: device inputs are used in place, host
// inputs are uploaded on the given stream.
struct DevIn {
    DBuf<double> buf;
    const double *p = nullptr;
    DevIn(const double *src, size_t n, bool dev, hipStream_t s)
    {
        if (dev) {
            p = src;
        } else {
            buf.upload(src, std::max<size_t>(n, 1), s);
            p = buf.p;
        }
    }
};

// Events that queued work may still wait on; destroyed by clear() once that
// work is known to have finished (or with the owner).
struct EventSet {
    std::vector<hipEvent_t> e;
    EventSet() = default;
    explicit EventSet(int count) { make(count); }
    EventSet(const EventSet &) = delete;
    EventSet &operator=(const EventSet &) = delete;
    ~EventSet() { clear(); }
    void make(int k)
    {
        clear();
        e.assign(k, nullptr);
        for (auto &x : e) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    }
    void clear()
    {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
        e.clear();
    }
};

struct NearTask {
    int32_t s0, s1, leaf, sf;   // source range [s0,s1) of leaf node `leaf`; surface
};
struct FarTask {
    int32_t g, i0, i1, sf;      // items [i0,i1) of target group g; surface
};

}  // namespace

struct tabi_ws {
    tabi::WsModel *m = nullptr;
};

struct tabi_ctx {
    tabi_params prm{};
    int32_t n = 0, ngroups = 0, p = 3, nw = 0, nk = 0;
    hipStream_t s = nullptr;
    HostTree tree;
    // geometry, tree order
    DBuf<double> tx, ty, tz, tnx, tny, tnz, tarea;
    DBuf<double> src8;                   // {x,y,z,nx,ny,nz,A*x_phi,A*x_psi} per face
    DBuf<float> src32;                   // mixed mode: same, float, leaf-local coordinates
    DBuf<float> mom_w32;                 // mixed mode: contraction vectors in float
    // nodes
    DBuf<double4> ngeo;                  // {xmid,ymid,zmid,radius}
    DBuf<int32_t> nnum, nchild, ncoff, nkids, nibeg;
    // interaction lists (per target group, CSR)
    std::vector<int32_t> h_far_off, h_near_off;
    DBuf<int32_t> far_off, near_off, far_node, near_leaf;
    DBuf<u64> far_mask, near_mask;
    DBuf<int32_t> cnt_far, cnt_near, need_mom, err_flag;
    // moments
    int32_t nmom = 0, nchunk = 0;
    DBuf<int32_t> node2mom, chunk_node, chunk_beg, chunk_end, mom_chunk_off, mom_node;
    DBuf<double> mom_part, mom_w;
    // tasks
    int32_t n_near_task = 0, n_far_task = 0;
    DBuf<NearTask> near_tasks;
    DBuf<FarTask> far_tasks;
    DBuf<int32_t> near_tgt, near_ipos, node_toff, node_nsplit, far_task_off;
    DBuf<double2> part_near, part_far;
    // preconditioner
    int32_t nleaf = 0, mmax = 0;
    std::vector<int64_t> h_blk_off;
    std::vector<int32_t> h_blk_m;
    DBuf<int64_t> blk_off;
    DBuf<int32_t> blk_m, blk_ibeg, piv_off, piv, perm, leaf_of, lu_info, lorder, lu_swrec;
    DBuf<double> lu_linv;
    DBuf<double> lu;
    DBuf<float> lu32;                    // mixed precision: the factors stored in FP32
    // factorisation in flight: its tile lists and events (freed by precond_finish)
    DBuf<int4> lu_tl, lu_tlo;
    DBuf<int2> lu_sw, lu_ts;
    DBuf<int32_t> lu_tab;
    DBuf<int64_t> lu_toff;
    EventSet pre_ev;
    // vectors
    DBuf<double> vx, vy, red;
    // blocks of vectors
    DBuf<double> wbd, mom_partb, mom_wb, blk_red;
    DBuf<float> wbf, mom_wb32;
    DBuf<double2> part_nearb, part_farb;
    // per-kernel timing (events on stream s)
    bool prof = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double kt_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t kt_n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // packed surfaces (one context over several surfaces): per surface the
    // first face (multiple of 64), face count, first node and first group,
    // each with an end entry; empty for a context of one surface
    std::vector<int32_t> sf_face0, sf_n, sf_node0, sf_group0;
    DBuf<int32_t> groot, gend;           // per group: root node, end of its surface's faces
    DBuf<int32_t> blk_sf;                // per leaf block: its surface
    DBuf<int32_t> group_sf, chunk_sf, mom_sf;   // per target group / moment chunk / moment node
    // surfaces whose field and preconditioner work is done (device flags, one
    // per surface; null: all): far, near and leaf-solve tasks of a surface
    // whose flag is 0 exit at once, leaving its entries of the outputs stale
    const int32_t *gate = nullptr;
    int nsurf() const { return sf_n.empty() ? 1 : (int)sf_n.size(); }
    int32_t nreal() const
    {
        if (sf_n.empty()) return n;
        int64_t t = 0;
        for (int32_t v : sf_n) t += v;
        return (int32_t)t;
    }
    // stats
    int64_t p_near = 0, p_far = 0, p_mom = 0, n_far_items = 0, n_near_items = 0;
    // a second stream for the near field, which runs beside the moments and
    // the far field of a product (taken from the stream pool on first use;
    // declared last, so it is synchronised before any buffer is freed)
    struct AuxStream {
        hipStream_t s = nullptr;
        hipEvent_t fork = nullptr, join = nullptr;
        AuxStream() = default;
        AuxStream(const AuxStream &) = delete;
        AuxStream &operator=(const AuxStream &) = delete;
        void ensure()
        {
            if (!fork) HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            if (!join) HIPCHK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
            if (!s) s = StreamPool::acquire();
        }
        ~AuxStream()
        {
            StreamPool::release(s);
            if (fork) (void)hipEventDestroy(fork);
            if (join) (void)hipEventDestroy(join);
        }
    } aux;
};

namespace {

// Brackets a launch sequence with events when profiling is on; the elapsed
// time is added to slot k once the stream reaches the end event.
struct KTimer {
    tabi_ctx *c;
    int k;
    KTimer(tabi_ctx *c_, int k_) : c(c_), k(k_)
    {
        if (c->prof) HIPCHK(hipEventRecord(c->ev[0], c->s));
    }
    ~KTimer() noexcept(false)
    {
        if (!c->prof) return;
        HIPCHK(hipEventRecord(c->ev[1], c->s));
        HIPCHK(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->kt_ms[k] += ms;
        c->kt_n[k] += 1;
    }
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// ---------------------------------------------------------------------------
// interaction lists: one wave per group of 64 consecutive targets walks the tree
// depth-first (children in order, as COMPP_TREE does, src/treecode3d_pb.f:359-436)
// carrying the mask of lanes that have failed the MAC at every ancestor.  At a
// node the MAC is evaluated per lane exactly as the reference evaluates it
// (src/treecode3d_pb.f:379-393): accepted lanes get a far item, the remaining
// lanes get a near item at a leaf or descend.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool mac_accept(double4 g, int numpar, double px, double py,
                                           double pz, double theta)
{
#pragma clang fp contract(off)
    const double tx = g.x - px, ty = g.y - py, tz = g.z - pz;
    const double dist = sqrt(tx * tx + ty * ty + tz * tz);
    return (g.w < dist * theta) && (numpar > 40);
}

__global__ void __launch_bounds__(256) k_lists(int pass, int n, int ngroups,
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    const double4 *__restrict__ ngeo, const int32_t *__restrict__ nnum,
    const int32_t *__restrict__ nchild, const int32_t *__restrict__ ncoff,
    const int32_t *__restrict__ nkids, double theta, int32_t *__restrict__ cnt_far,
    int32_t *__restrict__ cnt_near, const int32_t *__restrict__ far_off,
    const int32_t *__restrict__ near_off, int32_t *__restrict__ far_node,
    u64 *__restrict__ far_mask, int32_t *__restrict__ near_leaf, u64 *__restrict__ near_mask,
    int32_t *__restrict__ need_mom, int32_t *__restrict__ err_flag,
    const int32_t *__restrict__ groot, const int32_t *__restrict__ gend)
{
    __shared__ int32_t s_node[4][kStack];
    __shared__ u64 s_mask[4][kStack];
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int g = blockIdx.x * 4 + w;
    if (g >= ngroups) return;
    const int i = g * kWave + lane;
    // packed surfaces: the group's own tree and the end of its surface's faces
    const bool valid = i < (gend ? gend[g] : n);
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    int sp = 0;
    const u64 vmask = __ballot(valid);
    if (lane == 0) { s_node[w][0] = groot ? groot[g] : 0; s_mask[w][0] = vmask; }
    sp = 1;
    int nf = 0, nn = 0;
    int fbase = 0, nbase = 0;
    if (pass == 1) { fbase = far_off[g]; nbase = near_off[g]; }
    while (sp > 0) {
        --sp;
        const int nd = s_node[w][sp];
        const u64 m = s_mask[w][sp];
        const double4 geo = ngeo[nd];
        const int np = nnum[nd];
        bool acc = false;
        if ((m >> lane) & 1ull) acc = mac_accept(geo, np, px, py, pz, theta);
        const u64 accm = __ballot(acc) & m;
        if (accm) {
            if (pass == 1 && lane == 0) {
                far_node[fbase + nf] = nd;
                far_mask[fbase + nf] = accm;
                need_mom[nd] = 1;
            }
            ++nf;
        }
        const u64 rest = m & ~accm;
        if (rest) {
            const int nc = nchild[nd];
            if (nc == 0) {
                if (pass == 1 && lane == 0) {
                    near_leaf[nbase + nn] = nd;
                    near_mask[nbase + nn] = rest;
                }
                ++nn;
            } else {
                if (sp + nc > kStack) {
                    if (lane == 0) atomicOr(err_flag, 1);
                    break;
                }
                const int off = ncoff[nd];
                if (lane < nc) {
                    s_node[w][sp + nc - 1 - lane] = nkids[off + lane];
                    s_mask[w][sp + nc - 1 - lane] = rest;
                }
                sp += nc;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (pass == 0 && lane == 0) {
        cnt_far[g] = nf;
        cnt_near[g] = nn;
    }
}

// ---------------------------------------------------------------------------
// moments
// ---------------------------------------------------------------------------
template <int P>
struct MomShape {
    static constexpr int NK = (P + 1) * (P + 2) * (P + 3) / 6;   // |k| <= P
    static constexpr int L = P + 2;
    static constexpr int NJ = (L + 1) * (L + 2) * (L + 3) / 6;   // |j| <= P+2
    // length of each of a node's 8 far-field vectors: |k| <= P+1 (35 at p=3,
    // against NJ = 56 Taylor coefficients); see mom_final_body for the layout
    static constexpr int NW = (P + 2) * (P + 3) * (P + 4) / 6;
};

// inverse of tabi::tidx(): compact index -> exponents (i, j, k)
__device__ __forceinline__ void tdecode(int id, int &i, int &j, int &k)
{
    int d = 0;
    while ((d + 1) * (d + 2) * (d + 3) / 6 <= id) ++d;
    int r = id - d * (d + 1) * (d + 2) / 6;
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= r) ++s;
    k = r - s * (s + 1) / 2;
    j = s - k;
    i = d - s;
}

constexpr int kMomChunk = 512;

// raw moments of one chunk of a node's particles:
//   M_psi(k) = sum A_j psi_j d^k,  M_d(k) = sum (n_d A_j) phi_j d^k   (COMP_MS,
// src/treecode3d_pb.f:859-916, with pb_kernel's source charges, src/main.f90:671-692)
template <int P>
__device__ __forceinline__ void mom_raw_body(int c, int n, const int32_t *__restrict__ chunk_node,
    const int32_t *__restrict__ chunk_beg, const int32_t *__restrict__ chunk_end,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, const double *__restrict__ tarea,
    const double *__restrict__ x, double *__restrict__ part)
{
    constexpr int NK = MomShape<P>::NK;
    constexpr int NPART = 256 / NK;
    __shared__ double sd[7][kMomChunk];
    __shared__ double sred[4][256];
    const int t = threadIdx.x;
    const int nd = chunk_node[c], b = chunk_beg[c], e = chunk_end[c];
    const double4 geo = ngeo[nd];
    const int cnt = e - b;
    for (int q = t; q < cnt; q += 256) {
        const int i = b + q;
        const double a = tarea[i], phi = x[i], psi = x[n + i];
        sd[0][q] = tx[i] - geo.x;
        sd[1][q] = ty[i] - geo.y;
        sd[2][q] = tz[i] - geo.z;
        sd[3][q] = a * psi;
        sd[4][q] = tnx[i] * a * phi;
        sd[5][q] = tny[i] * a * phi;
        sd[6][q] = tnz[i] * a * phi;
    }
    __syncthreads();
    const int k = t % NK, part_id = t / NK;
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (part_id < NPART) {
        int e1, e2, e3;
        tdecode(k, e1, e2, e3);
        for (int q = part_id; q < cnt; q += NPART) {
            double pxv = 1.0, pyv = 1.0, pzv = 1.0;
            const double dx = sd[0][q], dy = sd[1][q], dz = sd[2][q];
            // powers by a fixed number of multiplications (by 1.0 past the
            // exponent: exact), so the lanes with different exponents do not
            // diverge
#pragma unroll
            for (int r = 0; r < P; ++r) {
                pxv *= r < e1 ? dx : 1.0;
                pyv *= r < e2 ? dy : 1.0;
                pzv *= r < e3 ? dz : 1.0;
            }
            const double mono = pxv * pyv * pzv;
            m0 += sd[3][q] * mono;
            m1 += sd[4][q] * mono;
            m2 += sd[5][q] * mono;
            m3 += sd[6][q] * mono;
        }
    }
    sred[0][t] = m0; sred[1][t] = m1; sred[2][t] = m2; sred[3][t] = m3;
    __syncthreads();
    if (t < NK) {
        double r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        for (int q = 0; q < NPART; ++q) {
            r0 += sred[0][q * NK + t]; r1 += sred[1][q * NK + t];
            r2 += sred[2][q * NK + t]; r3 += sred[3][q * NK + t];
        }
        double *o = part + (size_t)c * 4 * NK;
        o[t] = r0; o[NK + t] = r1; o[2 * NK + t] = r2; o[3 * NK + t] = r3;
    }
}

template <int P>
__global__ void __launch_bounds__(256) k_mom_raw(int n, const int32_t *__restrict__ chunk_node,
    const int32_t *__restrict__ chunk_beg, const int32_t *__restrict__ chunk_end,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, const double *__restrict__ tarea,
    const double *__restrict__ x, double *__restrict__ part, const int32_t *__restrict__ gate,
    const int32_t *__restrict__ chunk_sf)
{
    if (gate && !gate[chunk_sf[blockIdx.x]]) return;   // surface not being solved
    mom_raw_body<P>(blockIdx.x, n, chunk_node, chunk_beg, chunk_end, ngeo, tx, ty, tz, tnx, tny,
                    tnz, tarea, x, part);
}

// the same for column blockIdx.y of a block: vector act[col] at x + act[col]*xs,
// partials at part + col*pstride
template <int P>
__global__ void __launch_bounds__(256) k_mom_rawb(int n, const int32_t *__restrict__ chunk_node,
    const int32_t *__restrict__ chunk_beg, const int32_t *__restrict__ chunk_end,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, const double *__restrict__ tarea,
    const double *__restrict__ x, int64_t xs, const int32_t *__restrict__ act,
    double *__restrict__ part, int64_t pstride)
{
    const int col = blockIdx.y;
    mom_raw_body<P>(blockIdx.x, n, chunk_node, chunk_beg, chunk_end, ngeo, tx, ty, tz, tnx, tny,
                    tnz, tarea, x + (size_t)act[col] * xs, part + (size_t)col * pstride);
}

// Sum a node's chunk partials in chunk order and fold the moments with the
// derivative factors der_cof (src/main.f90:380-399) and the kernel combination
// of COMPP_TREE_PB (src/treecode3d_pb.f:479-485) into 8 vectors over |j| <= P+2:
//   far potential   peng1 = sum_j a0_j W[0][j] + ak_j W[1][j]
//   far normal der. peng2 = sum_d n0_d sum_j a0_j W[2+d][j] + ak_j W[5+d][j]
// where a0/ak are the Taylor coefficients for kappa = 0 and kappa.
// Only entries that can be nonzero are stored, NW = |{k : |k| <= P+1}| per
// vector: W[0], W[1] vanish at |j| = P+2 (stored at k = j), and W[2+d], W[5+d]
// wherever exponent d of j is 0 (stored at k = j - e_d).  The contractions
// read each vector contiguously and skip no nonzero term.
template <int P>
__device__ __forceinline__ void mom_final_body(int mi, const int32_t *__restrict__ mom_chunk_off,
    const double *__restrict__ part, double eps, double *__restrict__ mom_w,
    float *__restrict__ mom_w32 = nullptr)
{
    constexpr int NK = MomShape<P>::NK;
    constexpr int NJ = MomShape<P>::NJ;
    __shared__ double M[4][NK];
    const int t = threadIdx.x;
    const int c0 = mom_chunk_off[mi], c1 = mom_chunk_off[mi + 1];
    for (int q = t; q < 4 * NK; q += 256) {
        double s = 0.0;
        for (int c = c0; c < c1; ++c) s += part[(size_t)c * 4 * NK + q];
        M[q / NK][q % NK] = s;
    }
    __syncthreads();
    constexpr int NW = MomShape<P>::NW;
    double *W = mom_w + (size_t)mi * 8 * NW;
    float *W32 = mom_w32 ? mom_w32 + (size_t)mi * 8 * NW : nullptr;
    for (int j = t; j < NJ; j += 256) {
        int jv[3];
        tdecode(j, jv[0], jv[1], jv[2]);
        const int dj = jv[0] + jv[1] + jv[2];
        double ha = 0.0, hb = 0.0, hc[3] = {0, 0, 0}, he[3] = {0, 0, 0};
        if (dj <= P) ha = kOneOver4Pi * M[0][j];
        if (dj >= 1 && dj - 1 <= P) {
            for (int d = 0; d < 3; ++d) {
                if (jv[d] < 1) continue;
                int kv[3] = {jv[0], jv[1], jv[2]};
                kv[d] -= 1;
                const int kid = tabi::tidx(kv[0], kv[1], kv[2]);
                const double D = (double)(kv[d] + 1) * kOneOver4Pi;
                hb += D * M[1 + d][kid];
                hc[d] = D * M[0][kid];
            }
        }
        if (dj >= 2 && dj - 2 <= P) {
            for (int e = 0; e < 3; ++e) {
                double s = 0.0;
                for (int d = 0; d < 3; ++d) {
                    int kv[3] = {jv[0], jv[1], jv[2]};
                    kv[d] -= 1;
                    kv[e] -= 1;
                    if (kv[0] < 0 || kv[1] < 0 || kv[2] < 0) continue;
                    const int kid = tabi::tidx(kv[0], kv[1], kv[2]);
                    const double D = (d == e) ? (double)((kv[d] + 1) * (kv[d] + 2)) * kOneOver4Pi
                                              : (double)((kv[d] + 1) * (kv[e] + 1)) * kOneOver4Pi;
                    s += D * (-M[1 + d][kid]);
                }
                he[e] = s;
            }
        }
        double wv[8];
        wv[0] = ha - hb;
        wv[1] = -ha + eps * hb;
        for (int d = 0; d < 3; ++d) {
            wv[2 + d] = -hc[d] - he[d];
            wv[5 + d] = hc[d] / eps + he[d];
        }
        // the FP32 far field's copy, rounded as a separate conversion would
        if (dj <= P + 1) {
            for (int v = 0; v < 2; ++v) {
                W[v * NW + j] = wv[v];
                if (W32) W32[v * NW + j] = (float)wv[v];
            }
        }
        for (int d = 0; d < 3; ++d) {
            if (jv[d] < 1) continue;
            const int k = tabi::tidx(jv[0] - (d == 0), jv[1] - (d == 1), jv[2] - (d == 2));
            for (int v = 2 + d; v < 8; v += 3) {
                W[v * NW + k] = wv[v];
                if (W32) W32[v * NW + k] = (float)wv[v];
            }
        }
    }
}

template <int P>
__global__ void __launch_bounds__(256) k_mom_final(const int32_t *__restrict__ mom_chunk_off,
    const double *__restrict__ part, double eps, double *__restrict__ mom_w,
    float *__restrict__ mom_w32, const int32_t *__restrict__ gate, const int32_t *__restrict__ mom_sf)
{
    if (gate && !gate[mom_sf[blockIdx.x]]) return;     // surface not being solved
    mom_final_body<P>(blockIdx.x, mom_chunk_off, part, eps, mom_w, mom_w32);
}

// block column blockIdx.y: partials at part + col*pstride, vectors at
// mom_w + col*wstride
template <int P>
__global__ void __launch_bounds__(256) k_mom_finalb(const int32_t *__restrict__ mom_chunk_off,
    const double *__restrict__ part, int64_t pstride, double eps, double *__restrict__ mom_w,
    int64_t wstride)
{
    const int col = blockIdx.y;
    mom_final_body<P>(blockIdx.x, mom_chunk_off, part + (size_t)col * pstride, eps,
                      mom_w + (size_t)col * wstride);
}

// ---------------------------------------------------------------------------
// far field: one wave per task (a run of far items of one target group)
// ---------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(64) k_far(const FarTask *__restrict__ tasks, int ntask, int n,
    const int32_t *__restrict__ far_node, const u64 *__restrict__ far_mask,
    const int32_t *__restrict__ node2mom, const double *__restrict__ mom_w,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double kappa, double2 *__restrict__ part,
    const int32_t *__restrict__ gate)
{
    constexpr int L = MomShape<P>::L;
    constexpr int NJ = MomShape<P>::NJ;
    constexpr int NW = MomShape<P>::NW;
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const FarTask tk = tasks[task];
    if (gate && !gate[tk.sf]) return;           // surface not being solved
    const int lane = lane_id();
    const int i = tk.g * kWave + lane;
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    double acc1 = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    for (int it = tk.i0; it < tk.i1; ++it) {
        const int nd = far_node[it];
        const u64 m = far_mask[it];
        if (!((m >> lane) & 1ull)) continue;
        const double4 geo = ngeo[nd];
        const double *W = mom_w + (size_t)node2mom[nd] * 8 * NW;
        const double dx = px - geo.x, dy = py - geo.y, dz = pz - geo.z;
        {
            double a[NJ];
            tabi::taylor<L, false, double>(dx, dy, dz, 0.0, a);
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                acc1 = fma(a[q], W[q], acc1);
                ax = fma(a[tabi::tshift(q, 0)], W[2 * NW + q], ax);
                ay = fma(a[tabi::tshift(q, 1)], W[3 * NW + q], ay);
                az = fma(a[tabi::tshift(q, 2)], W[4 * NW + q], az);
            }
        }
        {
            double a[NJ];
            tabi::taylor<L, true, double>(dx, dy, dz, kappa, a);
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                acc1 = fma(a[q], W[NW + q], acc1);
                ax = fma(a[tabi::tshift(q, 0)], W[5 * NW + q], ax);
                ay = fma(a[tabi::tshift(q, 1)], W[6 * NW + q], ay);
                az = fma(a[tabi::tshift(q, 2)], W[7 * NW + q], az);
            }
        }
    }
    double p2 = 0.0;
    if (valid) p2 = tnx[i] * ax + tny[i] * ay + tnz[i] * az;
    part[(size_t)task * kWave + lane] = make_double2(acc1, p2);
}

// ---------------------------------------------------------------------------
// near field: one wave per task (one leaf source range for one target group)
// pair kernel of COMPP_DIRECT_PB, src/treecode3d_pb.f:965-1001; the self term
// is excluded (src/treecode3d_pb.f:1169-1172)
// ---------------------------------------------------------------------------
// 1/sqrt(r2) to full double precision: hardware estimate + two Newton steps
__device__ __forceinline__ double rsq_nr(double r2)
{
    double y = __builtin_amdgcn_rsq(r2);
    const double hx = 0.5 * r2;
    double t = hx * y;
    y = y * fma(-t, y, 1.5);
    t = hx * y;
    y = y * fma(-t, y, 1.5);
    return y;
}

// per-matvec source weights in the source records: {.., A*phi, A*dphi/dn}
__global__ void k_src_weights(int n, const double *__restrict__ tarea,
                              const double *__restrict__ x, double *__restrict__ src8,
                              const int32_t *__restrict__ gate = nullptr,
                              const int32_t *__restrict__ group_sf = nullptr)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (gate && !gate[group_sf[j / kWave]]) return;    // surface not being solved
    const double a = tarea[j];
    src8[(size_t)j * 8 + 6] = a * x[j];
    src8[(size_t)j * 8 + 7] = a * x[n + j];
}

// L1..L4 of one (target, source) pair; (dx,dy,dz) = source - target
struct PairL {
    double l1, l2, l3, l4;
};

__device__ __forceinline__ PairL pair_l(double dx, double dy, double dz, double r2,
                                        double vx, double vy, double vz, double v0x,
                                        double v0y, double v0z, double kappa, double kappa2,
                                        double eps, double inv_eps)
{
    const double inv = rsq_nr(r2);
    const double rs = r2 * inv;
    const double krs = kappa * rs;
    const double ek = exp(-krs);
    const double g0 = kOneOver4Pi * inv;
    const double gk = ek * g0;
    const double ct = (vx * dx + vy * dy + vz * dz) * inv;
    const double ct0 = (v0x * dx + v0y * dy + v0z * dz) * inv;
    const double tp1 = g0 * inv;
    const double tp2 = fma(krs, ek, ek);                 // (1 + kappa r) e^{-kappa r}
    const double g1 = ct * tp1;
    const double g10 = ct0 * tp1;
    const double g3 = (v0x * vx + v0y * vy + v0z * vz - 3.0 * ct0 * ct) * inv * tp1;
    PairL o;
    o.l1 = g1 * fma(-eps, tp2, 1.0);                     // G1 - eps G2
    o.l2 = g0 - gk;                                      // G0 - Gk
    o.l3 = fma(g3, tp2 - 1.0, -kappa2 * ct0 * ct * gk);   // G4 - G3
    o.l4 = g10 * fma(-inv_eps, tp2, 1.0);                // G10 - G20/eps
    return o;
}

__global__ void __launch_bounds__(64) k_near(const NearTask *__restrict__ tasks, 
    const int32_t *__restrict__ near_tgt, int ntask, int n,
    const double *__restrict__ src8, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double kappa, double eps, double2 *__restrict__ part,
    const int32_t *__restrict__ gate)
{
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const NearTask tk = tasks[task];
    if (gate && !gate[tk.sf]) return;           // surface not being solved
    const int lane = lane_id();
    const int i = near_tgt[(size_t)task * kWave + lane];   // this lane's target, -1: none
    const bool act = i >= 0;
    const bool valid = act;
    const double sx = valid ? tx[i] : 0.0, sy = valid ? ty[i] : 0.0, sz = valid ? tz[i] : 0.0;
    const double v0x = valid ? tnx[i] : 0.0, v0y = valid ? tny[i] : 0.0,
                 v0z = valid ? tnz[i] : 0.0;
    const double kappa2 = kappa * kappa, inv_eps = 1.0 / eps;
    double p1 = 0.0, p2 = 0.0;
    if (act) {
#pragma unroll 2
        for (int j = tk.s0; j < tk.s1; ++j) {
            const double4 ra = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8);
            const double4 rb = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8 + 4);
            const bool use = (j != i);
            const double dx = ra.x - sx, dy = ra.y - sy, dz = ra.z - sz;
            double r2 = dx * dx + dy * dy + dz * dz;
            r2 = use ? r2 : 1.0;
            const PairL L = pair_l(dx, dy, dz, r2, ra.w, rb.x, rb.y, v0x, v0y, v0z, kappa,
                                   kappa2, eps, inv_eps);
            const double w1 = use ? rb.z : 0.0, w2 = use ? rb.w : 0.0;
            p1 = fma(L.l1, w1, fma(L.l2, w2, p1));
            p2 = fma(L.l3, w1, fma(L.l4, w2, p2));
        }
    }
    part[(size_t)task * kWave + lane] = make_double2(p1, p2);
}

// ---------------------------------------------------------------------------
// mixed precision (precision = 1): FP32 pair arithmetic and Taylor recurrences,
// Kahan-compensated per-target sums inside a task, FP64 across tasks
// ---------------------------------------------------------------------------
struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// geometry part of the FP32 source records: coordinates relative to the
// centre of the source's leaf box (so nearby pairs keep their digits)
// sum of the set bits of n masks (an integer sum: exact in any order)
__global__ void __launch_bounds__(256) k_popc_sum(int64_t n, const u64 *__restrict__ mask,
                                                  unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long red[256];
    unsigned long long c = 0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        c += (unsigned long long)__popcll(mask[k]);
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// fl[i] = the leaf node holding particle i: one workgroup per leaf
__global__ void k_leaf_fill(const int32_t *__restrict__ leaves, const int32_t *__restrict__ nibeg,
                            const int32_t *__restrict__ nnum, int32_t *__restrict__ fl)
{
    const int nd = leaves[blockIdx.x];
    const int b = nibeg[nd], e = b + nnum[nd];
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) fl[i] = nd;
}

__global__ void k_src32_geom(int n, const int32_t *__restrict__ face_leafnode,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, float *__restrict__ src32)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double4 g = ngeo[face_leafnode[j]];
    float *r = src32 + (size_t)j * 8;
    r[0] = (float)(tx[j] - g.x);
    r[1] = (float)(ty[j] - g.y);
    r[2] = (float)(tz[j] - g.z);
    r[3] = (float)tnx[j];
    r[4] = (float)tny[j];
    r[5] = (float)tnz[j];
}

__global__ void k_src32_weights(int n, const double *__restrict__ tarea,
                                const double *__restrict__ x, float *__restrict__ src32,
                                const int32_t *__restrict__ gate, const int32_t *__restrict__ group_sf)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (gate && !gate[group_sf[j / kWave]]) return;    // surface not being solved
    const double a = tarea[j];
    src32[(size_t)j * 8 + 6] = (float)(a * x[j]);
    src32[(size_t)j * 8 + 7] = (float)(a * x[n + j]);
}

// FP32 near-field terms of one (target, source) pair (the pair kernel of
// COMPP_DIRECT_PB, src/treecode3d_pb.f:965-1001, as in the FP64 kernel): source
// record ra = {x, y, z, nx}, rb = {ny, nz, w1, w2}; target at (sx, sy, sz)
// with normal v0.  For the self pair (use false) 1/r is set to zero, which
// zeroes every term.
struct Near32 {
    float sx, sy, sz, v0x, v0y, v0z, kappa, kappa2, eps, inv_eps;
    __device__ __forceinline__ void terms(const float4 ra, const float4 rb, bool use,
                                          float &l1, float &l2, float &l3, float &l4) const
    {
        const float c4 = (float)kOneOver4Pi;
        const float dx = ra.x - sx, dy = ra.y - sy, dz = ra.z - sz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float inv = use ? __builtin_amdgcn_rsqf(r2) : 0.f;
        const float rs = r2 * inv;
        const float krs = kappa * rs;
        const float ek = __expf(-krs);
        const float g0 = c4 * inv;
        const float gk = ek * g0;
        const float ct = (ra.w * dx + rb.x * dy + rb.y * dz) * inv;
        const float ct0 = (v0x * dx + v0y * dy + v0z * dz) * inv;
        const float tp1 = g0 * inv;
        const float tp2 = fmaf(krs, ek, ek);
        const float g1 = ct * tp1, g10 = ct0 * tp1;
        const float g3 = (v0x * ra.w + v0y * rb.x + v0z * rb.y - 3.f * ct0 * ct) * inv * tp1;
        l1 = g1 * fmaf(-eps, tp2, 1.f);
        l2 = g0 - gk;
        l3 = fmaf(g3, tp2 - 1.f, -kappa2 * ct0 * ct * gk);
        l4 = g10 * fmaf(-inv_eps, tp2, 1.f);
    }
};

// Sources are taken kNearBlk at a time: their contributions are summed with
// FMAs into a block total, which then enters the compensated (Kahan) sum.
constexpr int kNearBlk = 4;

__global__ void __launch_bounds__(64) k_near32(const NearTask *__restrict__ tasks, 
    const int32_t *__restrict__ near_tgt, int ntask,
    int n, const float *__restrict__ src32, const double4 *__restrict__ ngeo,
    const double *__restrict__ tx, const double *__restrict__ ty,
    const double *__restrict__ tz, const double *__restrict__ tnx,
    const double *__restrict__ tny, const double *__restrict__ tnz, float kappa, float eps,
    double2 *__restrict__ part,
    const int32_t *__restrict__ gate)
{
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const NearTask tk = tasks[task];
    if (gate && !gate[tk.sf]) return;           // surface not being solved
    const int lane = lane_id();
    const int i = near_tgt[(size_t)task * kWave + lane];   // this lane's target, -1: none
    const bool act = i >= 0;
    const double4 lg = ngeo[tk.leaf];
    Near32 nk;
    nk.sx = act ? (float)(tx[i] - lg.x) : 0.f;
    nk.sy = act ? (float)(ty[i] - lg.y) : 0.f;
    nk.sz = act ? (float)(tz[i] - lg.z) : 0.f;
    nk.v0x = act ? (float)tnx[i] : 0.f;
    nk.v0y = act ? (float)tny[i] : 0.f;
    nk.v0z = act ? (float)tnz[i] : 0.f;
    nk.kappa = kappa; nk.kappa2 = kappa * kappa; nk.eps = eps; nk.inv_eps = 1.f / eps;
    Kahan p1, p2;
    if (act) {
        for (int j0 = tk.s0; j0 < tk.s1; j0 += kNearBlk) {
            float b1 = 0.f, b2 = 0.f;
#pragma unroll
            for (int u = 0; u < kNearBlk; ++u) {
                const int j = j0 + u;
                if (j < tk.s1) {
                    const float4 ra = *reinterpret_cast<const float4 *>(src32 + (size_t)j * 8);
                    const float4 rb = *reinterpret_cast<const float4 *>(src32 + (size_t)j * 8 + 4);
                    float l1, l2, l3, l4;
                    nk.terms(ra, rb, j != i, l1, l2, l3, l4);
                    b1 = fmaf(l1, rb.z, fmaf(l2, rb.w, b1));
                    b2 = fmaf(l3, rb.z, fmaf(l4, rb.w, b2));
                }
            }
            p1.add(b1);
            p2.add(b2);
        }
    }
    part[(size_t)task * kWave + lane] = make_double2((double)p1.s, (double)p2.s);
}

template <int P>
__global__ void __launch_bounds__(64) k_far32(const FarTask *__restrict__ tasks, int ntask, int n,
    const int32_t *__restrict__ far_node, const u64 *__restrict__ far_mask,
    const int32_t *__restrict__ node2mom, const float *__restrict__ mom_w,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, float kappa, double2 *__restrict__ part,
    const int32_t *__restrict__ gate)
{
    constexpr int L = MomShape<P>::L;
    constexpr int NJ = MomShape<P>::NJ;
    constexpr int NW = MomShape<P>::NW;
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const FarTask tk = tasks[task];
    if (gate && !gate[tk.sf]) return;           // surface not being solved
    const int lane = lane_id();
    const int i = tk.g * kWave + lane;
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    const float nx = valid ? (float)tnx[i] : 0.f, ny = valid ? (float)tny[i] : 0.f,
                nz = valid ? (float)tnz[i] : 0.f;
    Kahan s1, s2;
    for (int it = tk.i0; it < tk.i1; ++it) {
        const int nd = far_node[it];
        const u64 m = far_mask[it];
        if (!((m >> lane) & 1ull)) continue;
        const double4 geo = ngeo[nd];
        const float *W = mom_w + (size_t)node2mom[nd] * 8 * NW;
        const float dx = (float)(px - geo.x), dy = (float)(py - geo.y), dz = (float)(pz - geo.z);
        float acc1 = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
        {
            float a[NJ];
            tabi::taylor<L, false, float>(dx, dy, dz, 0.f, a);
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                acc1 = fmaf(a[q], W[q], acc1);
                ax = fmaf(a[tabi::tshift(q, 0)], W[2 * NW + q], ax);
                ay = fmaf(a[tabi::tshift(q, 1)], W[3 * NW + q], ay);
                az = fmaf(a[tabi::tshift(q, 2)], W[4 * NW + q], az);
            }
        }
        {
            float a[NJ];
            tabi::taylor<L, true, float>(dx, dy, dz, kappa, a);
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                acc1 = fmaf(a[q], W[NW + q], acc1);
                ax = fmaf(a[tabi::tshift(q, 0)], W[5 * NW + q], ax);
                ay = fmaf(a[tabi::tshift(q, 1)], W[6 * NW + q], ay);
                az = fmaf(a[tabi::tshift(q, 2)], W[7 * NW + q], az);
            }
        }
        s1.add(acc1);
        s2.add(nx * ax + ny * ay + nz * az);
    }
    part[(size_t)task * kWave + lane] = make_double2((double)s1.s, (double)s2.s);
}

// grid-stride: the factors of many packed surfaces exceed 2^32 entries, more
// work-items than one launch dimension can address
__global__ void k_to_float(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) b[i] = (float)a[i];
}
inline unsigned to_float_grid(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20); }

// Near-field task layout.  Item k of target group g (targets g*64 + lane)
// lists source leaf nd for the lanes set in near_mask[k].  Leaf nd's targets
// are gathered in item order (k ascending, so groups ascending) and, inside an
// item, lanes ascending: item k's j-th listed target sits at position
// p = ipos[k] + j of nd's list.  The list is cut into waves of 64 and each wave
// into nsplit[nd] source runs, so position p is lane p % 64 of the tasks
// toff[nd] + (p / 64) * nsplit[nd] + s, s < nsplit[nd].
struct NearSlots {
    const int32_t *__restrict__ off;
    const int32_t *__restrict__ leaf;
    const u64 *__restrict__ mask;
    const int32_t *__restrict__ ipos;
    const int32_t *__restrict__ toff;
    const int32_t *__restrict__ nsplit;
};

// the target rows of the near tasks (padding lanes hold -1 from a prior fill):
// one wave per target group, lane = the target's lane in the group
__global__ void __launch_bounds__(64) k_near_tgt(NearSlots ns, int32_t *__restrict__ tgt)
{
    const int g = blockIdx.x, lane = threadIdx.x;
    const u64 below = (1ull << lane) - 1;
    for (int k = ns.off[g]; k < ns.off[g + 1]; ++k) {
        const u64 m = ns.mask[k];
        if (!((m >> lane) & 1)) continue;
        const int p = ns.ipos[k] + __popcll(m & below);
        const int nd = ns.leaf[k], nsp = ns.nsplit[nd];
        const int64_t t0 = (int64_t)ns.toff[nd] + (int64_t)(p / kWave) * nsp;
        for (int s = 0; s < nsp; ++s) tgt[(t0 + s) * kWave + p % kWave] = g * kWave + lane;
    }
}

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t *__restrict__ a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// y = pre*x - sum of the target's partials: its near-field tasks (items of
// its group in list order, then source runs) and then the far tasks of its
// group (TREE_COMPP_PB's combination, src/treecode3d_pb.f:1154-1155,1177-1178)
__global__ void k_reduce(int n, NearSlots ns, const int32_t *__restrict__ far_task_off,
    const double2 *__restrict__ part_near, const double2 *__restrict__ part_far,
    const double *__restrict__ x, double pre1, double pre2, double *__restrict__ y,
    const int32_t *__restrict__ gate, const int32_t *__restrict__ group_sf)
{
    // the group's item records are loaded by the wave's lanes together (one
    // round of dependent loads instead of one per item), then walked in order
    __shared__ u64 s_mask[4][kWave];
    __shared__ int s_ipos[4][kWave], s_toff[4][kWave], s_nsp[4][kWave];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = i / kWave, lane = i % kWave, w = threadIdx.x / kWave;
    if (g * kWave >= n) return;                  // the whole wave is past the end
    if (gate && !gate[group_sf[g]]) return;      // surface not being solved
    const u64 below = (1ull << lane) - 1;
    double p1 = 0.0, p2 = 0.0;
    const int k0 = ns.off[g], k1 = ns.off[g + 1];
    for (int kb = k0; kb < k1; kb += kWave) {
        const int cnt = min(kWave, k1 - kb);
        if (lane < cnt) {
            const int k = kb + lane, nd = ns.leaf[k];
            s_mask[w][lane] = ns.mask[k];
            s_ipos[w][lane] = ns.ipos[k];
            s_toff[w][lane] = ns.toff[nd];
            s_nsp[w][lane] = ns.nsplit[nd];
        }
        __builtin_amdgcn_wave_barrier();
        for (int kk = 0; kk < cnt; ++kk) {
            const u64 m = s_mask[w][kk];
            if (!((m >> lane) & 1)) continue;
            const int p = s_ipos[w][kk] + __popcll(m & below);
            const int nsp = s_nsp[w][kk];
            const double2 *pt = part_near + ((size_t)s_toff[w][kk] + (size_t)(p / kWave) * nsp) * kWave + p % kWave;
            for (int s = 0; s < nsp; ++s) {
                const double2 v = pt[(size_t)s * kWave];
                p1 += v.x; p2 += v.y;
            }
        }
        __builtin_amdgcn_wave_barrier();         // every lane has read the records
    }
    if (i >= n) return;
    for (int t = far_task_off[g]; t < far_task_off[g + 1]; ++t) {
        const double2 v = part_far[(size_t)t * kWave + lane];
        p1 += v.x; p2 += v.y;
    }
    y[i] = pre1 * x[i] - p1;
    y[n + i] = pre2 * x[n + i] - p2;
}

// ---------------------------------------------------------------------------
// leaf-block preconditioner
// ---------------------------------------------------------------------------
// Assemble every leaf block, column-major with leading dimension m = 2*nl, in
// the reference's layout (src/treecode3d_pb.f:1439-1528): one wave per source
// face j (four per workgroup), lanes over the targets i of j's leaf, so a
// leaf of ~150 faces keeps most lanes busy.  flags bit 0 selects
// the matvec-consistent placement of the L2/L3 couplings.
constexpr int kAsmPerWave = 4;   // source faces per wave in the block assembly
__global__ void __launch_bounds__(256) k_lu_assemble(int n, const int32_t *__restrict__ leaf_of,
    const int32_t *__restrict__ blk_ibeg, const int32_t *__restrict__ blk_m,
    const int64_t *__restrict__ blk_off, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, const double *__restrict__ tarea, double kappa,
    double eps, int flags, double *__restrict__ lu)
{
    const int lane = threadIdx.x & 63;
    const int jw = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kAsmPerWave;
    for (int j = jw; j < min(n, jw + kAsmPerWave); ++j) {
    const int l = leaf_of[j];
    if (l < 0) continue;                 // padding between packed surfaces
    const int ib = blk_ibeg[l], m = blk_m[l], nl = m / 2;
    const int jj = j - ib;
    double *A = lu + blk_off[l];
    const double inv_eps = 1.0 / eps;
    const double pre1 = 0.5 * (1.0 + eps), pre2 = 0.5 * (1.0 + inv_eps);
    const double rx = tx[j], ry = ty[j], rz = tz[j];
    const double vx = tnx[j], vy = tny[j], vz = tnz[j], aj = tarea[j];
    const bool swap23 = flags & 1;
    for (int ii = lane; ii < nl; ii += kWave) {
        const int i = ib + ii;
        double e1, e2, e3, e4;
        if (i == j) {
            A[(size_t)jj * m + ii] = pre1;
            A[(size_t)jj * m + ii + nl] = 0.0;
            A[(size_t)(jj + nl) * m + ii] = 0.0;
            A[(size_t)(jj + nl) * m + ii + nl] = pre2;
            continue;
        }
        const double dx = rx - tx[i], dy = ry - ty[i], dz = rz - tz[i];
        const double v0x = tnx[i], v0y = tny[i], v0z = tnz[i];
        const double rs = sqrt(dx * dx + dy * dy + dz * dz);
        const double irs = 1.0 / rs;
        const double g0 = kOneOver4Pi * irs;
        const double krs = kappa * rs;
        const double ek = exp(-krs);
        const double gk = ek * g0;
        const double ct = (vx * dx + vy * dy + vz * dz) * irs;
        const double ct0 = (v0x * dx + v0y * dy + v0z * dz) * irs;
        const double tp1 = g0 * irs;
        const double tp2 = (1.0 + krs) * ek;
        const double g10 = ct0 * tp1, g20 = tp2 * g10;
        const double g1 = ct * tp1, g2 = tp2 * g1;
        const double g3 = (v0x * vx + v0y * vy + v0z * vz - 3.0 * ct0 * ct) * irs * tp1;
        const double g4 = tp2 * g3 - kappa * kappa * ct0 * ct * gk;
        e1 = -(g1 - eps * g2) * aj;
        e2 = -(g0 - gk) * aj;
        e3 = -(g4 - g3) * aj;
        e4 = -(g10 - g20 * inv_eps) * aj;
        if (swap23) { const double t = e2; e2 = e3; e3 = t; }
        A[(size_t)jj * m + ii] = e1;
        A[(size_t)jj * m + ii + nl] = e2;
        A[(size_t)(jj + nl) * m + ii] = e3;
        A[(size_t)(jj + nl) * m + ii + nl] = e4;
    }
    }
}

// Batched LU: leaf blocks are processed largest first (lorder[] sorts leaves by
// block size, descending), so the blocks still active at panel step k0 are a
// prefix of lorder and every launch covers only active blocks.
//
// Two-level blocking.  Columns are factored in strips of kLuW (four panels of
// kLuNB).  Inside a strip the panels are right-looking over the strip's own
// columns only.  The columns beyond the strip take the row interchanges of
// every panel as they come, but none of the strip's arithmetic until its end,
// so their rows always hold the original entries (interchanged) and the
// interchanges move consistent data.  When the strip is done, its rows of
// those columns are solved at once, U12 = L11^-1 A12 with the 128x128 unit
// lower triangle, and everything below the strip is updated by one rank-kLuW
// product.  The factors are DGETRF's (the unblocked steps of DGETF2
// within each panel, src/dgetrf_dep.f:757-914); only the order in which
// the trailing products are accumulated differs.
constexpr int kLuW = 4 * kLuNB;   // strip width

// Interchange record of one panel (pivot rows piv[c] of the rows k0 + c,
// c < nb, applied in order): the touched rows dst[q], q < nr -- the panel
// rows, then each distinct pivot row below the panel in order of first use --
// and src[q], the row whose contents end up in dst[q].  Built by one wave
// with cross-lane exchanges (lane c holds pivot c; lane q tracks slot q);
// writes dst and src to the arrays given and returns nr in every lane.
constexpr int kSwRec = 4 * kLuNB + 4;   // ints per block: dst[64], src[64], nr
__device__ __forceinline__ int lu_swap_record(int lane, int k0, int nbe, int pr, int *dst,
                                              int *src)
{
    // first occurrence of this lane's pivot row
    int first = lane;
    for (int c2 = 0; c2 < kLuNB; ++c2) {
        const int o = __shfl(pr, c2);
        if (c2 < first && c2 < lane && o == pr) first = c2;
    }
    const bool isnew = lane < nbe && pr >= k0 + nbe && first == lane;
    const u64 newmask = __ballot(isnew);
    const int nr = nbe + __popcll(newmask);
    // slot of this lane's pivot row in the touched-row list
    int slot = lane;
    if (lane < nbe)
        slot = pr < k0 + nbe ? pr - k0 : nbe + __popcll(newmask & ((1ull << first) - 1));
    // the list itself, held one entry per lane: panel rows, then new extras
    int row = lane < nbe ? k0 + lane : 0;
    {
        // the extra with rank r sits in the lane of the r-th set bit of newmask
        const int r = lane - nbe;
        int idx = 0;
        if (r >= 0 && lane < nr) {
            u64 mm = newmask;
            for (int k = 0; k < r; ++k) mm &= mm - 1;
            idx = __ffsll((long long)mm) - 1;
        }
        const int o = __shfl(pr, idx);
        if (r >= 0 && lane < nr) row = o;
    }
    // position tracking: after the swaps (c, slot[c]) in order, slot q holds
    // the original contents of slot p
    int p = lane;
    for (int c = 0; c < nbe; ++c) {
        const int sl = __shfl(slot, c);
        const int pc = __shfl(p, c), ps = __shfl(p, sl);
        if (lane == c) p = ps;
        else if (lane == sl) p = pc;
    }
    const int srow = __shfl(row, p);
    if (lane < nr) { dst[lane] = row; src[lane] = srow; }
    return nr;
}

// The double held by lane `lane` (a compile-time or wave-uniform index) of
// the wave, read into every lane through two 32-bit lane reads.
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A double / an int from another lane of the same row of 16 lanes (DPP
// control Ctrl: quad permutations and row rotations only, so every source
// lane exists).  Every lane of the wave must be active.
template <int Ctrl>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const unsigned lo2 = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, Ctrl, 0xF, 0xF, false);
    const unsigned hi2 = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, Ctrl, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi2 << 32) | lo2));
}
template <int Ctrl>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, Ctrl, 0xF, 0xF, false);
}

// Maximum (minimum) over the wave, in every lane: quad_perm [1,0,3,2] and
// [2,3,0,1], row_ror 4 and 8 leave each row of 16 lanes holding its row's
// extreme, then the four rows' values are read from lanes 0, 16, 32, 48.  No
// LDS round trips: the arg-max of a pivoting panel column is on its critical
// path.  Every lane of the wave must be active.
__device__ __forceinline__ double wave_max_d(double v)
{
    v = fmax(v, dpp_d<0xB1>(v));
    v = fmax(v, dpp_d<0x4E>(v));
    v = fmax(v, dpp_d<0x124>(v));
    v = fmax(v, dpp_d<0x128>(v));
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}
__device__ __forceinline__ int wave_min_i(int v)
{
    v = min(v, dpp_i<0xB1>(v));
    v = min(v, dpp_i<0x4E>(v));
    v = min(v, dpp_i<0x124>(v));
    v = min(v, dpp_i<0x128>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Panel factorisation with partial pivoting, for a panel in which the
// optimistic pass met a column whose diagonal entry is not its first maximum
// in magnitude: the panel is reloaded and factored with the first-maximum rule
// (IDAMAX), one barrier per column.  Per column every wave finds its
// candidate (largest magnitude, first row among equals) with DPP reductions
// and the candidate's lane leaves the value, the row index and the whole row in
// LDS, beside the row at the diagonal; after the barrier every thread combines
// the waves' candidates in the same order, the two rows trade places through
// those copies and the rank-1 update reads the pivot row from its LDS copy.
// The per-column buffers alternate by column parity: a thread writes column
// c + 2's only after the barrier of column c + 1, which every thread passes
// only after its reads of column c.  s_cr holds 2 * NW * kLuNB doubles.  Out
// of line, so the common pass keeps its registers.
template <int RPT, int kPanelThreads>
__device__ __attribute__((noinline)) void lu_panel_pivoted(int k0, int m, int nbe,
    double *__restrict__ A, float *__restrict__ A32, int32_t *__restrict__ pv,
    int32_t *__restrict__ info_l,
    double (*s_best)[kPanelThreads / kWave], int (*s_bi)[kPanelThreads / kWave],
    double (*s_row)[2][kLuNB], double *__restrict__ s_cr, int32_t *s_piv)
{
    constexpr int NW = kPanelThreads / kWave;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    double p[RPT][kLuNB];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int row = k0 + t + j * kPanelThreads;
#pragma unroll
        for (int c = 0; c < kLuNB; ++c)
            p[j][c] = (row < m && c < nbe) ? A[(size_t)(k0 + c) * m + row] : 0.0;
    }
    int bad = 0;
#pragma clang loop unroll(full)
    for (int c = 0; c < kLuNB; ++c) {
        if (c < nbe) {  // nbe is uniform over the workgroup
        const int col = k0 + c, pb = c & 1;
        // this thread's candidate among its rows col..m-1 (rows grow with j)
        double best = -1.0;
        int bi = 0x7fffffff, jb = 0;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int row = k0 + t + j * kPanelThreads;
            if (row >= col && row < m) {
                const double v = fabs(p[j][c]);
                if (v > best) { best = v; bi = row; jb = j; }
            }
        }
        // the wave's: largest magnitude, then the first row holding it (with
        // one row per thread rows grow with the lane)
        const double wmax = wave_max_d(best);
        int wrow;
        if (RPT == 1) {
            const u64 hit = __ballot(best == wmax);
            wrow = __builtin_amdgcn_readlane(bi, __ffsll((long long)hit) - 1);
        } else {
            wrow = wave_min_i(best == wmax ? bi : 0x7fffffff);
        }
        double *cr = s_cr + (size_t)(pb * NW + w) * kLuNB;
        if (wmax < 0.0) {                        // no row of this wave is a candidate
            if (lane == 0) { s_best[pb][w] = -1.0; s_bi[pb][w] = 0x7fffffff; }
        } else if (best == wmax && bi == wrow) { // exactly one lane: rows are distinct
            s_best[pb][w] = wmax;
            s_bi[pb][w] = wrow;
#pragma unroll
            for (int j = 0; j < RPT; ++j)
                if (j == jb) {
#pragma unroll
                    for (int cc = 0; cc < kLuNB; ++cc) cr[cc] = p[j][cc];
                }
        }
        if (t == c) {                            // row col is row 0 of thread c
#pragma unroll
            for (int cc = 0; cc < kLuNB; ++cc) s_row[pb][0][cc] = p[0][cc];
        }
        __syncthreads();
        // every thread combines the waves' candidates (same order everywhere)
        double gmax = s_best[pb][0];
        int prow = s_bi[pb][0], qw = 0;
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            const double v = s_best[pb][q];
            const int r = s_bi[pb][q];
            if (v > gmax || (v == gmax && r < prow)) { gmax = v; prow = r; qw = q; }
        }
        const double *pr = s_cr + (size_t)(pb * NW + qw) * kLuNB;   // the pivot row
        if (t == 0) { s_piv[c] = prow; pv[col] = prow; }
        if (prow != col) {
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int row = k0 + t + j * kPanelThreads;
                if (row == prow) {
#pragma unroll
                    for (int cc = 0; cc < kLuNB; ++cc) p[j][cc] = s_row[pb][0][cc];
                } else if (row == col) {
#pragma unroll
                    for (int cc = 0; cc < kLuNB; ++cc) p[j][cc] = pr[cc];
                }
            }
        }
        const double pivot = pr[c];
        if (pivot == 0.0) {
            bad = col + 1;
        } else {
            const double rinv = 1.0 / pivot;
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int row = k0 + t + j * kPanelThreads;
                if (row > col && row < m) {
                    p[j][c] *= rinv;
                    const double lv = p[j][c];
#pragma unroll
                    for (int cc = c + 1; cc < kLuNB; ++cc) p[j][cc] = fma(-lv, pr[cc], p[j][cc]);
                }
            }
        }
        }
    }
    if (bad && t == 0) atomicCAS(info_l, 0, bad);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int row = k0 + t + j * kPanelThreads;
        if (row < m) {
#pragma unroll
            for (int c = 0; c < kLuNB; ++c)
                if (c < nbe) {
                    A[(size_t)(k0 + c) * m + row] = p[j][c];
                    if (A32) A32[(size_t)(k0 + c) * m + row] = (float)p[j][c];
                }
        }
    }
}

// Panel factorisation of columns k0..k0+nb-1: partial pivoting with the
// first-maximum rule (IDAMAX), reciprocal scaling and rank-1 updates restricted
// to the panel (the unblocked step of DGETRF2's semantics,
// src/dgetrf_dep.f:757-914).  One workgroup of kPanelThreads per block holds
// the panel in registers, RPT rows per thread (row k0 + t + j*kPanelThreads);
// each column step is a block-wide arg-max, a row exchange through LDS and a
// broadcast of the pivot row.  Afterwards the panel's interchanges are applied
// to the strip's earlier columns K0..k0-1 (staged in LDS).
// Row interchanges of panel k0 on the block's column tiles that
// k_lu_swap_trsm does not take (those left of the panel's 64-column tile and
// those wholly beyond the strip), by the panel's workgroup when the panel made
// interchanges: each tile's touched rows are read into registers, then, after
// a barrier, written to their destinations.  Out of line: only panels that
// pivot reach it.
// The same interchanges on the strip's earlier columns and the block's other
// columns together (every column but the panel's and those k_lu_swap_trsm
// takes), when at most kSwapCols rows are touched: one thread per column reads
// the touched rows' entries of its column, then writes them to their
// destinations (a permutation of the same rows), so no barrier is needed.
constexpr int kSwapCols = 16;
template <int kPanelThreads>
__device__ __attribute__((noinline)) void lu_swap_columns(int k0, int m, int nbe, double *A, float *A32,
                                                          int nr, const int *rows, const int *src)
{
    if (nr <= 0) return;                      // a zero pivot without interchanges
    const int K0 = k0 - k0 % kLuW, send = min(m, K0 + kLuW);
    const bool has_right = send > k0 + kLuNB;
    const int r0 = (k0 + kLuNB) / 64, r1 = (send - 1) / 64;
    int sr[kSwapCols], dr[kSwapCols];
#pragma unroll
    for (int q = 0; q < kSwapCols; ++q) {
        sr[q] = src[min(q, nr - 1)];
        dr[q] = rows[min(q, nr - 1)];
    }
    for (int j = threadIdx.x; j < m; j += kPanelThreads) {
        if (j >= k0 && j < k0 + nbe) continue;                          // the panel
        if (has_right && j / 64 >= r0 && j / 64 <= r1) continue;         // k_lu_swap_trsm's
        double *col = A + (size_t)j * m;
        double v[kSwapCols];
#pragma unroll
        for (int q = 0; q < kSwapCols; ++q)
            if (q < nr) v[q] = col[sr[q]];
#pragma unroll
        for (int q = 0; q < kSwapCols; ++q)
            if (q < nr) col[dr[q]] = v[q];
        if (A32) {
            float *col32 = A32 + (size_t)j * m;
#pragma unroll
            for (int q = 0; q < kSwapCols; ++q)
                if (q < nr) col32[dr[q]] = (float)v[q];
        }
    }
}

template <int kPanelThreads>
__device__ __attribute__((noinline)) void lu_swap_rest(int k0, int m, int nbe, double *A, float *A32,
                                                       int nr, const int *rows, const int *src)
{
    const int t = threadIdx.x;
    const int K0 = k0 - k0 % kLuW, send = min(m, K0 + kLuW);
    // the tiles [r0, r1] belong to k_lu_swap_trsm (none when send <= k0 + kLuNB)
    const bool has_right = send > k0 + kLuNB;
    const int r0 = (k0 + kLuNB) / 64, r1 = (send - 1) / 64;
    constexpr int kTs = 64 * 64 / kPanelThreads;
    const int nct = (m + 63) / 64;
    for (int ct = 0; ct < nct; ++ct) {
        if (has_right && ct >= r0 && ct <= r1) continue;   // uniform
        const int j0 = ct * 64;
        double v[kTs];
#pragma unroll
        for (int i = 0; i < kTs; ++i) {
            const int e = t + kPanelThreads * i, q = e % 64, j = min(j0 + e / 64, m - 1);
            v[i] = A[(size_t)j * m + src[min(q, nr - 1)]];
        }
        __syncthreads();                         // every touched row is read
#pragma unroll
        for (int i = 0; i < kTs; ++i) {
            const int e = t + kPanelThreads * i, q = e % 64, j = j0 + e / 64;
            if (q < nr && j < m && (j < K0 || j >= k0 + nbe)) {
                A[(size_t)j * m + rows[q]] = v[i];
                if (A32) A32[(size_t)j * m + rows[q]] = (float)v[i];
            }
        }
        __syncthreads();
    }
}

// At least three waves per SIMD: left to itself the compiler gives the 256-
// and 512-thread forms 248 registers (two waves per SIMD); capped at 168,
// more panels are in flight (measured on the 484k-face surface: set-up
// 19.9 -> 19.6 ms; four waves per SIMD, 128 registers, was slower).
template <int RPT, int kPanelThreads>
__global__ void __launch_bounds__(kPanelThreads, 3) k_lu_panel(int k0, int first,
    const int32_t *__restrict__ lorder, const int32_t *__restrict__ blk_m,
    const int64_t *__restrict__ blk_off, const int32_t *__restrict__ piv_off,
    double *__restrict__ lu, int32_t *__restrict__ piv, int32_t *__restrict__ info,
    int32_t *__restrict__ swrec, float *__restrict__ lu32)
{
    constexpr int NW = kPanelThreads / kWave;
    // per column: the waves' maxima and rows, and the two exchanged rows;
    // double-buffered by column parity, so a column step needs two barriers
    __shared__ double s_best[2][NW];
    __shared__ int s_bi[2][NW];
    __shared__ double s_row[2][2][kLuNB];
    __shared__ int s_prow;
    __shared__ int32_t s_piv[kLuNB];
    __shared__ int s_dst[2 * kLuNB], s_src[2 * kLuNB];
    const int l = lorder[first + blockIdx.x], t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m = blk_m[l];
    const int nbe = min(kLuNB, m - k0);
    double *A = lu + blk_off[l];
    // the FP32 copy of the factors (mixed precision), written as entries are
    // finished: the panel's columns here, their later interchanges below
    float *A32 = lu32 ? lu32 + blk_off[l] : nullptr;
    int32_t *pv = piv + piv_off[l];
    double p[RPT][kLuNB];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int row = k0 + t + j * kPanelThreads;
#pragma unroll
        for (int c = 0; c < kLuNB; ++c)
            p[j][c] = (row < m && c < nbe) ? A[(size_t)(k0 + c) * m + row] : 0.0;
    }
    // Optimistic pass: each column's diagonal entry as its pivot.  A thread
    // notes when one of its rows below the diagonal is strictly larger in
    // magnitude (IDAMAX takes the first maximum, so a tie keeps the diagonal)
    // or the pivot is zero.  If no thread did, partial pivoting would have
    // chosen the diagonal in every column and the factors are exactly those of
    // the pivoted pass below (the same operations); otherwise the panel is
    // reloaded and factored with partial pivoting.
    // Wave 0 holds the panel's diagonal block (rows k0..k0+31 are its lanes
    // 0..31) and factors its 64 rows column by column, the pivot row taken from
    // lane c with cross-lane reads, without barriers; it then leaves the
    // finished pivot rows (U11) and their reciprocals in LDS, and after one
    // barrier every other row applies them in the same order (row-wise
    // substitution: the same multiplications and fused updates as the
    // column-by-column elimination, so the factors are bit-identical to it).
    int need = 0;
    __shared__ double s_u[kLuNB][kLuNB + 1];
    __shared__ double s_rinv[kLuNB];
    if (w == 0) {
#pragma unroll
        for (int c = 0; c < kLuNB; ++c) {
            if (c < nbe) {
                const int col = k0 + c;
                const double pivot = readlane_d(p[0][c], c);
                const double ap = fabs(pivot), rinv = 1.0 / pivot;
                need |= pivot == 0.0;
                if (lane == 0) s_rinv[c] = rinv;
                double u[kLuNB];
#pragma unroll
                for (int cc = c + 1; cc < kLuNB; ++cc) u[cc] = readlane_d(p[0][cc], c);
                const int row = k0 + t;
                if (row > col && row < m) {
                    need |= fabs(p[0][c]) > ap;
                    p[0][c] *= rinv;
                    const double lv = p[0][c];
#pragma unroll
                    for (int cc = c + 1; cc < kLuNB; ++cc) p[0][cc] = fma(-lv, u[cc], p[0][cc]);
                }
            }
        }
        if (t < nbe) {
#pragma unroll
            for (int cc = 0; cc < kLuNB; ++cc) s_u[t][cc] = p[0][cc];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int row = k0 + t + j * kPanelThreads;
        if ((j > 0 || w > 0) && row < m) {          // every such row is below the panel
#pragma unroll
            for (int c = 0; c < kLuNB; ++c) {
                if (c < nbe) {
                    need |= fabs(p[j][c]) > fabs(s_u[c][c]);
                    p[j][c] *= s_rinv[c];
                    const double lv = p[j][c];
#pragma unroll
                    for (int cc = c + 1; cc < kLuNB; ++cc)
                        p[j][cc] = fma(-lv, s_u[c][cc], p[j][cc]);
                }
            }
        }
    }
    const bool fast = !__syncthreads_or(need);
    if (fast) {
        if (t < nbe) { s_piv[t] = k0 + t; pv[k0 + t] = k0 + t; }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int row = k0 + t + j * kPanelThreads;
        if (row < m) {
#pragma unroll
            for (int c = 0; c < kLuNB; ++c)
                if (c < nbe) {
                    A[(size_t)(k0 + c) * m + row] = p[j][c];
                    if (A32) A32[(size_t)(k0 + c) * m + row] = (float)p[j][c];
                }
        }
    }
    } else {
        static_assert(2 * NW <= kLuNB + 1, "candidate rows exceed s_u");
        lu_panel_pivoted<RPT, kPanelThreads>(k0, m, nbe, A, A32, pv, info + l, s_best, s_bi, s_row,
                                             &s_u[0][0], s_piv);
    }
    __syncthreads();
    // interchange record for the other columns (read by k_lu_swap_trsm)
    int32_t *rec = swrec + (size_t)l * kSwRec;
    if (w == 0) {
        const int nr = lu_swap_record(lane, k0, nbe, lane < nbe ? s_piv[lane] : -1, s_dst, s_src);
        if (lane < nr) { rec[lane] = s_dst[lane]; rec[2 * kLuNB + lane] = s_src[lane]; }
        if (lane == 0) { rec[4 * kLuNB] = nr; rec[4 * kLuNB + 1] = fast; s_prow = nr; }
    }
    // the panel's interchanges on the strip's earlier columns K0..k0-1: every
    // touched row is read, then written to its destination
    const int K0 = k0 - k0 % kLuW, ncl = k0 - K0;
    if (fast) return;
    __syncthreads();
    const int nr = s_prow;
    if (nr <= kSwapCols) {
        // few touched rows: every column at once, one thread per column
        lu_swap_columns<kPanelThreads>(k0, m, nbe, A, A32, nr, s_dst, s_src);
        return;
    }
    if (ncl > 0) {
        // element e = t + kPanelThreads i: row slot e % 64, column e / 64
        constexpr int kEs = (2 * kLuNB * (kLuW - kLuNB) + kPanelThreads - 1) / kPanelThreads;
        double v[kEs];
#pragma unroll
        for (int i = 0; i < kEs; ++i) {
            const int e = t + kPanelThreads * i, q = e % 64, jj = min(e / 64, ncl - 1);
            v[i] = A[(size_t)(K0 + jj) * m + s_src[min(q, nr - 1)]];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kEs; ++i) {
            const int e = t + kPanelThreads * i, q = e % 64, jj = e / 64;
            if (q < nr && jj < ncl) {
                A[(size_t)(K0 + jj) * m + s_dst[q]] = v[i];
                if (A32) A32[(size_t)(K0 + jj) * m + s_dst[q]] = (float)v[i];
            }
        }
        __syncthreads();
    }
    // and on the block's column tiles outside the strip that k_lu_swap_trsm
    // does not take
    lu_swap_rest<kPanelThreads>(k0, m, nbe, A, A32, nr, s_dst, s_src);
}

// Inverse of the unit lower triangle L11 of panel k0 (rows and columns
// k0..k0+nb-1, zero-padded to kLuNB) for every active block, so that
// U12 = L11^-1 A12 becomes an independent product per entry.  One wave per
// block; lane c < kLuNB forms column c by forward substitution.  The inverse
// is stored column-major in the block's slot for the panel's place in its
// strip (kLuW / kLuNB slots per block), where the strip's final solve finds
// all four.
__global__ void __launch_bounds__(64) k_lu_linv(int k0, int first,
    const int32_t *__restrict__ lorder, const int32_t *__restrict__ blk_m,
    const int64_t *__restrict__ blk_off, const double *__restrict__ lu,
    double *__restrict__ linv)
{
    __shared__ double Ls[kLuNB][kLuNB + 1];
    __shared__ double Xs[kLuNB][kLuNB + 1];
    const int l = lorder[first + blockIdx.x], t = threadIdx.x;
    const int m = blk_m[l];
    const int nbe = min(kLuNB, m - k0);
    const double *A = lu + blk_off[l];
    double v[kLuNB * kLuNB / 64];
#pragma unroll
    for (int i = 0; i < kLuNB * kLuNB / 64; ++i) {
        const int e = t + 64 * i, r = e % kLuNB, c = e / kLuNB;
        v[i] = A[(size_t)(k0 + min(c, nbe - 1)) * m + k0 + min(r, nbe - 1)];
    }
#pragma unroll
    for (int i = 0; i < kLuNB * kLuNB / 64; ++i) {
        const int e = t + 64 * i, r = e % kLuNB, c = e / kLuNB;
        Ls[r][c] = (c < r && r < nbe) ? v[i] : 0.0;
        Xs[r][c] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
    if (t < kLuNB) {
        // the column in registers; L's zero padding (rows past nb) leaves the
        // identity entries unchanged, so the full 32-row sweep computes what
        // the nb-row one does, with every product independent of LDS writes
        double x[kLuNB];
#pragma unroll
        for (int r = 0; r < kLuNB; ++r) x[r] = Xs[r][t];
#pragma unroll
        for (int k = 0; k < kLuNB - 1; ++k) {
#pragma unroll
            for (int r = k + 1; r < kLuNB; ++r) x[r] = fma(-Ls[r][k], x[k], x[r]);
        }
#pragma unroll
        for (int r = 0; r < kLuNB; ++r) Xs[r][t] = x[r];
    }
    __syncthreads();
    double *Li = linv + ((size_t)l * (kLuW / kLuNB) + (k0 % kLuW) / kLuNB) * kLuNB * kLuNB;
#pragma unroll
    for (int i = 0; i < kLuNB * kLuNB / 64; ++i) {
        const int e = t + 64 * i;
        Li[e] = Xs[e % kLuNB][e / kLuNB];
    }
}

// Row interchanges of panel k0 on every column outside the strip's factored
// part, and U12 = L11^-1 A12 for the strip's columns right of the panel (their
// rows are up to date: inside the strip the factorisation is right-looking).
// A workgroup takes one of the block's 64-column tiles that hold strip
// columns right of the panel (the panel kernel takes the other tiles); the
// touched rows (the panel rows and the pivot rows below them) are staged in
// LDS with coalesced loads.  The tile's columns outside the strip take the
// interchanges only.
__global__ void __launch_bounds__(256) k_lu_swap_trsm(int k0, const int2 *__restrict__ tiles,
    const int32_t *__restrict__ blk_m, const int64_t *__restrict__ blk_off,
    double *__restrict__ lu, const int32_t *__restrict__ swrec, const double *__restrict__ linv,
    float *__restrict__ lu32)
{
    __shared__ double Li[kLuNB][kLuNB + 1];          // L11^-1 (row-major)
    __shared__ double T[2 * kLuNB][64 + 1];
    __shared__ int rows[2 * kLuNB], src[2 * kLuNB];
    __shared__ int s_nr;
    const int2 tl = tiles[blockIdx.x];           // {block, column tile}
    const int l = tl.x;
    if (l < 0) return;                         // padding of the XCD-grouped list
    const int m = blk_m[l];
    const int nbe = min(kLuNB, m - k0);
    const int K0 = k0 - k0 % kLuW, send = min(m, K0 + kLuW);
    const int j0 = tl.y * 64;
    double *A = lu + blk_off[l];
    const int32_t *rec = swrec + (size_t)l * kSwRec;
    const int t = threadIdx.x;
    if (t < 2 * kLuNB) { rows[t] = rec[t]; src[t] = rec[2 * kLuNB + t]; }
    if (t == 0) s_nr = rec[4 * kLuNB];
    // the tile holds strip columns right of the panel
    const bool right = j0 + 63 >= k0 + nbe && j0 < send;
    if (!right && rec[4 * kLuNB + 1]) return;  // interchanges only, and the panel made none
    // staging loops have fixed trip counts and clamped addresses, so their
    // loads are issued together
    if (right) {
        const double *Lg = linv + ((size_t)l * (kLuW / kLuNB) + (k0 - K0) / kLuNB) * kLuNB * kLuNB;
        double v[kLuNB * kLuNB / 256];
#pragma unroll
        for (int i = 0; i < kLuNB * kLuNB / 256; ++i) v[i] = Lg[t + 256 * i];
#pragma unroll
        for (int i = 0; i < kLuNB * kLuNB / 256; ++i) {
            const int e = t + 256 * i;
            Li[e % kLuNB][e / kLuNB] = v[i];
        }
    }
    __syncthreads();
    const int nr = s_nr;
    auto col_ok = [&](int j) { return j < m && (j < K0 || j >= k0 + nbe); };
    // touched rows: element e = t + 256 i is row slot e % 64, column e / 64
    constexpr int kTs = 2 * kLuNB * 64 / 256;
    {
        double v[kTs];
#pragma unroll
        for (int i = 0; i < kTs; ++i) {
            const int e = t + 256 * i, q = e % 64, j = min(j0 + e / 64, m - 1);
            v[i] = A[(size_t)j * m + src[min(q, nr - 1)]];
        }
#pragma unroll
        for (int i = 0; i < kTs; ++i) {
            const int e = t + 256 * i, q = e % 64, jj = e / 64;
            if (q < nr) T[q][jj] = col_ok(j0 + jj) ? v[i] : 0.0;
        }
    }
    __syncthreads();    // the rows are staged in their post-interchange order
    // U12 = L11^-1 A12 for the strip columns right of the panel: thread
    // (column t % 64, row group t / 64) forms rows rg + 4i of its column
    if (right) {
        const int jc = t % 64, rg = t / 64, j = j0 + jc;
        const bool mine = j < send && j >= k0 + nbe;
        double acc[kLuNB / 4];
#pragma unroll
        for (int i = 0; i < kLuNB / 4; ++i) acc[i] = 0.0;
        if (mine) {
#pragma unroll 4
            for (int c = 0; c < nbe; ++c) {
                const double tc = T[c][jc];
#pragma unroll
                for (int i = 0; i < kLuNB / 4; ++i) acc[i] = fma(Li[rg + 4 * i][c], tc, acc[i]);
            }
        }
        __syncthreads();
        if (mine) {
#pragma unroll
            for (int i = 0; i < kLuNB / 4; ++i)
                if (rg + 4 * i < nbe) T[rg + 4 * i][jc] = acc[i];
        }
    }
    __syncthreads();
    // the panel rows of the strip's columns are final here (U12), so is every
    // interchanged entry left of the strip; the FP32 copy takes them all
    float *A32 = lu32 ? lu32 + blk_off[l] : nullptr;
#pragma unroll
    for (int i = 0; i < kTs; ++i) {
        const int e = t + 256 * i, q = e % 64, j = j0 + e / 64;
        if (q < nr && col_ok(j)) {
            A[(size_t)j * m + rows[q]] = T[q][e / 64];
            if (A32) A32[(size_t)j * m + rows[q]] = (float)T[q][e / 64];
        }
    }
}


typedef double f64x4 __attribute__((ext_vector_type(4)));

// U12 = L11^-1 A12 for the columns beyond a finished strip K0..K0+kLuW-1,
// whose strip rows still hold their original (interchanged) entries.  A
// workgroup takes 64 columns of one block, staged in LDS column-major; wave w
// owns columns 16w..16w+15 and solves them panel by panel on the FP64 matrix
// cores, R = A_p - L[rows of p][K0..K0+32p) X and X_p = (L11_p)^-1 R with the
// panel inverses of k_lu_linv, so no wave reads another wave's columns.
// Operands as in k_lu_gemm: A L[row = lane%16][k = lane/16] straight from
// global memory (16 consecutive rows of 4 columns), B X[k = lane/16][col =
// lane%16] from LDS; result register v = D[lane/16 + 4v][lane%16].
constexpr int kStS = kLuW + 2;   // LDS column stride in doubles: conflict-free operand reads
__global__ void __launch_bounds__(256, 2) k_lu_strip_trsm(int K0, const int2 *__restrict__ tiles,
    const int32_t *__restrict__ blk_m, const int64_t *__restrict__ blk_off,
    double *__restrict__ lu, const double *__restrict__ linv, float *__restrict__ lu32)
{
    __shared__ double T[64 * kStS];
    const int2 tl = tiles[blockIdx.x];           // {block, column tile beyond the strip}
    const int l = tl.x;
    if (l < 0) return;                         // padding of the XCD-grouped list
    const int m = blk_m[l];
    const int j0 = K0 + kLuW + tl.y * 64;
    double *A = lu + blk_off[l];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, li = lane & 15, lk = lane >> 4;
    // stage rows K0..K0+kLuW-1 of the tile's columns: element e = t + 256 i is
    // row e % kLuW of column e / kLuW (columns past m are clamped, never stored)
    constexpr int kTs = kLuW * 64 / 256;
    {
        double v[kTs];
#pragma unroll
        for (int i = 0; i < kTs; ++i) {
            const int e = t + 256 * i;
            v[i] = A[(size_t)min(j0 + e / kLuW, m - 1) * m + K0 + e % kLuW];
        }
#pragma unroll
        for (int i = 0; i < kTs; ++i) {
            const int e = t + 256 * i;
            T[(e / kLuW) * kStS + e % kLuW] = v[i];
        }
    }
    __syncthreads();
    const double *L = A + (size_t)K0 * m + K0;  // L11, column-major, leading dimension m
    const double *Lv = linv + (size_t)l * (kLuW / kLuNB) * kLuNB * kLuNB;
    double *X = T + (16 * w + li) * kStS;       // this lane's column
    // a wave's LDS accesses complete in order, so the lanes of one wave read
    // each other's results below without a barrier
#pragma unroll
    for (int p = 0; p < kLuW / kLuNB; ++p) {
        // the panel's A operands are loaded together before its products, so
        // their latency is paid once per panel rather than once per product
        const double *Li = Lv + p * kLuNB * kLuNB;  // column-major 32x32
        double al[2][(kLuW - kLuNB) / 4], ai[2][kLuNB / 4];
#pragma unroll
        for (int q = 0; q < kLuNB * p / 4; ++q)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                al[h][q] = -L[(size_t)(4 * q + lk) * m + kLuNB * p + 16 * h + li];
#pragma unroll
        for (int q = 0; q < kLuNB / 4; ++q)
#pragma unroll
            for (int h = 0; h < 2; ++h) ai[h][q] = Li[(4 * q + lk) * kLuNB + 16 * h + li];
        f64x4 acc[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[h][v] = X[kLuNB * p + 16 * h + lk + 4 * v];
#pragma unroll
        for (int q = 0; q < kLuNB * p / 4; ++q) {
            const double b = X[4 * q + lk];
#pragma unroll
            for (int h = 0; h < 2; ++h)
                acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(al[h][q], b, acc[h], 0, 0, 0);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) X[kLuNB * p + 16 * h + lk + 4 * v] = acc[h][v];
        f64x4 x[2] = {f64x4{0.0, 0.0, 0.0, 0.0}, f64x4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int q = 0; q < kLuNB / 4; ++q) {
            const double b = X[kLuNB * p + 4 * q + lk];
#pragma unroll
            for (int h = 0; h < 2; ++h)
                x[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[h][q], b, x[h], 0, 0, 0);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) X[kLuNB * p + 16 * h + lk + 4 * v] = x[h][v];
    }
    __syncthreads();
    float *A32 = lu32 ? lu32 + blk_off[l] : nullptr;   // U12 beyond the strip is final
#pragma unroll
    for (int i = 0; i < kTs; ++i) {
        const int e = t + 256 * i, c = e / kLuW;
        if (j0 + c < m) {
            A[(size_t)(j0 + c) * m + K0 + e % kLuW] = T[c * kStS + e % kLuW];
            if (A32) A32[(size_t)(j0 + c) * m + K0 + e % kLuW] = (float)T[c * kStS + e % kLuW];
        }
    }
}

// Trailing update A22 -= L21 U12 on 64x64 tiles from a precomputed tile list
// {block, tile row, tile col, column end} on the FP64 matrix cores.  L21 holds
// the columns kb..kb+KD-1 (KD = kLuNB for the in-strip update after a panel,
// kLuW for the update below a finished strip), taken kLuNB at a time through
// LDS; rows and columns start at kb + KD; columns stop at the tile's column
// end.  Each of the four waves owns a 32x32 quadrant as 2x2 blocks of
// v_mfma_f64_16x16x4f64: A operand L21[row = lane%16][k = lane/16], B operand
// U12[k = lane/16][col = lane%16], result register v = D[lane/16 + 4v][lane%16].

template <int KD>
__global__ void __launch_bounds__(256, KD == kLuNB ? 1 : 3) k_lu_gemm(int kb, const int4 *__restrict__ tiles,
    const int32_t *__restrict__ blk_m, const int64_t *__restrict__ blk_off,
    double *__restrict__ lu)
{
    __shared__ double sL[kLuNB][64 + 1];   // [k][row]
    __shared__ double sU[kLuNB][64 + 1];   // [k][column]
    const int4 tl = tiles[blockIdx.x];
    const int l = tl.x;
    if (l < 0) return;                     // padding of the XCD-grouped list
    const int m = blk_m[l];
    const int cend = tl.w;
    const int r0 = kb + KD + tl.y * 64, c0 = kb + KD + tl.z * 64;
    double *A = lu + blk_off[l];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int rw = (w & 1) * 32, cw = (w >> 1) * 32, li = lane & 15, lk = lane >> 4;
    // after a panel (one chunk) the trailing entries are loaded before the
    // product so their latency overlaps the arithmetic; below a strip (four
    // chunks) they are loaded after it, which frees their registers for a
    // third wave per SIMD
    constexpr bool kEarlyC = KD == kLuNB;
    double cv[2][2][4];
    auto load_c = [&]() {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int col = c0 + cw + 16 * bj + li;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = r0 + rw + 16 * bi + lk + 4 * v;
                    cv[bi][bj][v] = (row < m && col < cend) ? A[(size_t)col * m + row] : 0.0;
                }
            }
    };
    if (kEarlyC) load_c();
    f64x4 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = f64x4{0.0, 0.0, 0.0, 0.0};
    // chunk staging: element e = t + 256 i of L21 (column e / 64, row e % 64)
    // and of U12 (column e / kLuNB, row e % kLuNB); rows and columns past the
    // tile's end are clamped (their products land in outputs that are not
    // stored), so every load is unconditional and a chunk's 16 loads are in
    // flight together -- the next chunk's while this one is multiplied
    constexpr int kSt = kLuNB * 64 / 256;
    double lst[kSt], ust[kSt];
    auto fetch = [&](int k1) {
#pragma unroll
        for (int i = 0; i < kSt; ++i) {
            const int e = t + 256 * i;
            const int row = min(r0 + (e & 63), m - 1);
            lst[i] = A[(size_t)(k1 + (e >> 6)) * m + row];
            const int col = min(c0 + e / kLuNB, cend - 1);
            ust[i] = A[(size_t)col * m + k1 + e % kLuNB];
        }
    };
    fetch(kb);
#pragma unroll 1
    for (int k1 = kb; k1 < kb + KD; k1 += kLuNB) {
        if (k1 > kb) __syncthreads();
#pragma unroll
        for (int i = 0; i < kSt; ++i) {
            const int e = t + 256 * i;
            sL[e >> 6][e & 63] = lst[i];
            sU[e % kLuNB][e / kLuNB] = ust[i];
        }
        __syncthreads();
        if (k1 + kLuNB < kb + KD) fetch(k1 + kLuNB);
#pragma unroll
        for (int kk = 0; kk < kLuNB; kk += 4) {
            const double a0 = sL[kk + lk][rw + li], a1 = sL[kk + lk][rw + 16 + li];
            const double b0 = sU[kk + lk][cw + li], b1 = sU[kk + lk][cw + 16 + li];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
    if (!kEarlyC) load_c();
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int col = c0 + cw + 16 * bj + li;
            double *Ac = A + (size_t)col * m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r0 + rw + 16 * bi + lk + 4 * v;
                if (row < m && col < cend) Ac[row] = cv[bi][bj][v] - acc[bi][bj][v];
            }
        }
}

// Composite row permutation of each leaf's pivot sequence: (P r)[k] = r[perm[k]]
// (the interchanges of DGETRS/DLASWP applied in order, src/dgetrf_dep.f:346-448).
// One wave per leaf: the identity is written by all lanes, and only a block
// with an interchange goes through the sequential pass (lane 0), since the
// interchanges compose in order.
__global__ void __launch_bounds__(kWave) k_lu_perm(const int32_t *__restrict__ blk_m,
                          const int32_t *__restrict__ piv_off, const int32_t *__restrict__ piv,
                          int32_t *__restrict__ perm)
{
    const int l = blockIdx.x, lane = threadIdx.x;
    const int m = blk_m[l];
    const int32_t *pv = piv + piv_off[l];
    int32_t *pm = perm + piv_off[l];
    int swaps = 0;
    for (int k = lane; k < m; k += kWave) {
        pm[k] = k;
        swaps |= pv[k] != k;
    }
    if (!__syncthreads_or(swaps)) return;        // the block is one wave
    if (lane != 0) return;
    for (int k = 0; k < m; ++k) {
        const int p = pv[k];
        if (p != k) { const int32_t t = pm[k]; pm[k] = pm[p]; pm[p] = t; }
    }
}

// z = M^-1 r for every leaf block with the cached factors: permuted gather, then
// blocked forward (unit lower) and back (upper) substitution.  Each 64x64
// diagonal tile is held row-per-lane in registers and solved with shuffles; the
// rows outside the tile are updated 64 at a time, each lane issuing the tile's
// 64 column loads back to back so the factor stream keeps many loads in
// flight.  z lives in LDS.
constexpr int kSolveB = 64;
constexpr int kSolveNW = 3;      // waves per leaf block in k_lu_solve
constexpr int kHalf = 32;
constexpr int kDiagH = 32;       // diagonal-tile factors in flight per pass

// s - sum_{c < bs} A[c*m] * zt[c] for one row, loads issued 32 at a time
// A factor entry that is used only when `use` holds.  Float factors are
// loaded unconditionally from a clamped, in-block address and zeroed after the
// widening: a guarded float load becomes a branch around load and conversion,
// which serialises the loads (a wait per load); guarded double loads stay in
// flight and are cheaper than the clamped addresses.
template <typename FT>
__device__ __forceinline__ FT factor_ld(const FT *__restrict__ base, size_t off, size_t off_clamped,
                                        bool use)
{
    if constexpr (sizeof(FT) == 8) {
        return use ? base[off] : FT(0);
    } else {
        const FT v = base[off_clamped];
        return use ? v : FT(0);
    }
}

template <typename FT>
__device__ __forceinline__ double gemv_row(const FT *__restrict__ Ac, int m, int bs,
                                           const double *zt, double s)
{
    // entries in flight per pass (all 64 for float factors took the kernel to
    // 256 registers and tripled its time)
    constexpr int H = kHalf;
    double s1 = 0.0;
#pragma unroll 1
    for (int h = 0; h < kSolveB; h += H) {
        // the loads first, all in flight, then the widening (converting inside
        // the guarded load made the compiler wait for each load in turn)
        FT af[H];
#pragma unroll
        for (int c = 0; c < H; ++c)
            af[c] = factor_ld(Ac, (size_t)(h + c) * m, (size_t)min(h + c, bs - 1) * m, h + c < bs);
#pragma unroll
        for (int c = 0; c < H; c += 2) {
            s = fma(-(double)af[c], (h + c < bs) ? zt[h + c] : 0.0, s);
            s1 = fma(-(double)af[c + 1], (h + c + 1 < bs) ? zt[h + c + 1] : 0.0, s1);
        }
    }
    return s + s1;
}

// NW waves per leaf, one barrier per 64-row tile, with one tile of
// lookahead.  Forward step bi starts with z of tile bi-1 final: wave 0 applies
// it to tile bi's rows and solves the 64x64 diagonal tile bi (the solved entry
// of each step reaches the other lanes through a scalar-register lane read,
// not an LDS shuffle: this chain of 64 dependent steps per tile is the
// critical path of a large leaf), while the other waves apply z of tile bi-1
// to the rows below tile bi, 64 rows per chunk taken from an LDS counter;
// wave 0 joins them once its tile is solved.  So the bulk of the factor stream
// keeps flowing while the diagonal tile is solved, instead of waiting for it.
// The back substitution runs the same way upwards.  Each row still takes the
// tiles' contributions in the same order with the same arithmetic, so z is
// bit-identical to the tile-by-tile schedule.  Leaves are visited in the order
// lorder (largest first) so the longest solves start first.
//
// FT is the stored factor type: double, or float for the mixed-precision form
// (factors computed in FP64 and rounded once for storage; the substitutions
// accumulate in FP64), which halves the stream.
template <int NW, typename FT>
__global__ void __launch_bounds__(64 * NW) k_lu_solve(int n, const int32_t *__restrict__ lorder,
    const int32_t *__restrict__ blk_ibeg, const int32_t *__restrict__ blk_m,
    const int64_t *__restrict__ blk_off, const int32_t *__restrict__ piv_off,
    const FT *__restrict__ lu, const int32_t *__restrict__ perm, const double *__restrict__ r,
    double *__restrict__ z, const int32_t *__restrict__ gate, const int32_t *__restrict__ blk_sf)
{
    extern __shared__ double zs[];
    __shared__ int s_next[2];                  // chunk counters, alternating by step
    const int l = lorder[blockIdx.x], t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (gate && !gate[blk_sf[l]]) return;      // surface not being solved
    const int m = blk_m[l], nl = m / 2, ib = blk_ibeg[l];
    const FT *A = lu + blk_off[l];
    const int32_t *pm = perm + piv_off[l];
    if (t == 0) { s_next[0] = 0; s_next[1] = 0; }
    for (int k = t; k < m; k += 64 * NW) {
        const int src = pm[k];
        zs[k] = src < nl ? r[ib + src] : r[n + ib + src - nl];
    }
    __syncthreads();
    const int nb = (m + kSolveB - 1) / kSolveB;
    // rows [rbeg, rend) take the contribution of the tile at column ks (width
    // bs), 64 rows per chunk, chunks handed out by the counter
    auto bulk = [&](int *next, int rbeg, int rend, int ks, int bs) {
        const int nch = rend > rbeg ? (rend - rbeg + 63) / 64 : 0;
        for (;;) {
            int q = 0;
            if (lane == 0) q = atomicAdd(next, 1);
            q = __builtin_amdgcn_readlane(q, 0);
            if (q >= nch) break;
            const int row = rbeg + 64 * q + lane;
            if (row < rend) zs[row] = gemv_row(A + (size_t)ks * m + row, m, bs, zs + ks, zs[row]);
        }
    };
    int step = 0;
    for (int bi = 0; bi < nb; ++bi, ++step) {
        const int kb = bi * kSolveB, bs = min(kSolveB, m - kb);
        if (t == 0) s_next[(step + 1) & 1] = 0;
        if (w == 0) {
            const bool in = lane < bs;
            double zl = in ? zs[kb + lane] : 0.0;
            if (bi > 0 && in)                  // tile bi-1's contribution to tile bi
                zl = gemv_row(A + (size_t)(kb - kSolveB) * m + kb + lane, m, kSolveB, zs + kb - kSolveB, zl);
#pragma unroll 1
            for (int h = 0; h < kSolveB; h += kDiagH) {
                FT tf[kDiagH];
#pragma unroll
                for (int c = 0; c < kDiagH; ++c)
                    tf[c] = factor_ld(A, (size_t)(kb + h + c) * m + kb + lane,
                                      (size_t)(kb + min(h + c, bs - 1)) * m + kb + min(lane, bs - 1),
                                      in && h + c < lane);
                double tr[kDiagH];
#pragma unroll
                for (int c = 0; c < kDiagH; ++c) tr[c] = (double)tf[c];
#pragma unroll
                for (int c = 0; c < kDiagH; ++c) zl -= tr[c] * readlane_d(zl, h + c);
            }
            if (in) zs[kb + lane] = zl;
        }
        if (bi > 0) bulk(&s_next[step & 1], kb + kSolveB, m, kb - kSolveB, kSolveB);
        __syncthreads();
    }
    for (int bi = nb - 1; bi >= 0; --bi, ++step) {
        const int kb = bi * kSolveB, bs = min(kSolveB, m - kb);
        const int kb1 = kb + kSolveB, bs1 = min(kSolveB, m - kb1);   // tile bi+1
        if (t == 0) s_next[(step + 1) & 1] = 0;
        if (w == 0) {
            const bool in = lane < bs;
            double zl = in ? zs[kb + lane] : 0.0;
            if (bi < nb - 1 && in)             // tile bi+1's contribution to tile bi
                zl = gemv_row(A + (size_t)kb1 * m + kb + lane, m, bs1, zs + kb1, zl);
#pragma unroll 1
            for (int h = kSolveB - kDiagH; h >= 0; h -= kDiagH) {
                FT tf[kDiagH];
#pragma unroll
                for (int c = 0; c < kDiagH; ++c)
                    tf[c] = factor_ld(A, (size_t)(kb + h + c) * m + kb + lane,
                                      (size_t)(kb + min(h + c, bs - 1)) * m + kb + min(lane, bs - 1),
                                      in && h + c >= lane && h + c < bs);
                double tr[kDiagH];
#pragma unroll
                for (int c = 0; c < kDiagH; ++c) tr[c] = (double)tf[c];
#pragma unroll
                for (int c = kDiagH - 1; c >= 0; --c) {
                    if (lane == h + c && in) zl = zl / tr[c];
                    const double zc = readlane_d(zl, h + c);
                    if (lane < h + c) zl -= tr[c] * zc;
                }
            }
            if (in) zs[kb + lane] = zl;
        }
        if (bi < nb - 1) bulk(&s_next[step & 1], 0, kb, kb1, bs1);
        __syncthreads();
    }
    for (int q = t; q < nl; q += 64 * NW) {
        z[ib + q] = zs[q];
        z[n + ib + q] = zs[nl + q];
    }
}

// ---------------------------------------------------------------------------
// right-hand side and energy
// ---------------------------------------------------------------------------
// Charge records {x, y, z, q} for the right-hand side.
__global__ void k_charge_rec(int nq, const double *__restrict__ qpos, const double *__restrict__ q,
                             double4 *__restrict__ rec)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nq) rec[k] = make_double4(qpos[3 * k], qpos[3 * k + 1], qpos[3 * k + 2], q[k]);
}

// b_i = sum_k q_k/(4 pi eps_p r), b_{n+i} = sum_k q_k cos/(4 pi eps_p r^2)
// (src/main.f90:597-607), one thread per face, the charge records read
// wave-uniformly (scalar loads, no LDS staging); the second half is written
// ho elements after the first (ho = n, or the packed length)
__global__ void __launch_bounds__(256) k_rhs(int n, int nq, const double4 *__restrict__ qrec,
    double eps_p, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double *__restrict__ b, int64_t ho)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    const double vx = valid ? tnx[i] : 0.0, vy = valid ? tny[i] : 0.0, vz = valid ? tnz[i] : 0.0;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll 4
    for (int k = 0; k < nq; ++k) {
        const double4 c = qrec[k];
        const double dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const double inv = rsq_nr(dx * dx + dy * dy + dz * dz);
        const double ct = (vx * dx + vy * dy + vz * dz) * inv;
        b1 += c.w * inv;
        b2 += c.w * ct * inv * inv;
    }
    if (valid) {
        const double s = 1.0 / (4.0 * kPi * eps_p);
        b[i] = b1 * s;
        b[ho + i] = b2 * s;
    }
}

// Mixed precision: the right-hand side with FP32 pair terms.  A charge record
// keeps its position in FP64 (the differences are taken in FP64, then
// rounded) and its charge in FP32; per face the terms of 16 charges are
// summed in FP32, those sums in FP64.  b1 += q/r, b2 += q (n.d)/r^3 as k_rhs
// forms them (d = charge - face).
struct QRec32 {
    double x, y, z;
    float q, pad;
};
__global__ void k_charge_rec32(int nq, const double *__restrict__ qpos, const double *__restrict__ q,
                               QRec32 *__restrict__ rec)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nq) rec[k] = QRec32{qpos[3 * k], qpos[3 * k + 1], qpos[3 * k + 2], (float)q[k], 0.f};
}

constexpr int kRhsSub = 16;      // charges per FP32 partial sum
__global__ void __launch_bounds__(256) k_rhs32(int n, int nq, const QRec32 *__restrict__ qrec,
    double eps_p, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double *__restrict__ b, int64_t ho)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    const float vx = valid ? (float)tnx[i] : 0.f, vy = valid ? (float)tny[i] : 0.f,
                vz = valid ? (float)tnz[i] : 0.f;
    double b1 = 0.0, b2 = 0.0;
    for (int k0 = 0; k0 < nq; k0 += kRhsSub) {
        const int k1 = min(nq, k0 + kRhsSub);
        float s1 = 0.f, s2 = 0.f;
#pragma unroll 4
        for (int k = k0; k < k1; ++k) {
            const QRec32 c = qrec[k];
            const float dx = (float)(c.x - px), dy = (float)(c.y - py), dz = (float)(c.z - pz);
            const float inv = __builtin_amdgcn_rsqf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
            const float qi = c.q * inv;
            const float nd = fmaf(vx, dx, fmaf(vy, dy, vz * dz));
            s1 += qi;
            s2 = fmaf(qi, nd * (inv * inv), s2);
        }
        b1 += (double)s1;
        b2 += (double)s2;
    }
    if (valid) {
        const double s = 1.0 / (4.0 * kPi * eps_p);
        b[i] = b1 * s;
        b[ho + i] = b2 * s;
    }
}

// Potential at each charge from the surface solution (potential_molecule,
// src/main.f90:627-655): tasks of (64 charges, a run of kEnergyChunk faces), one
// wave each, faces read wave-uniformly from the source records (which hold
// A*phi, A*dphi/dn of the solution); partials are summed in chunk order.
constexpr int kEnergyChunk = 2048;
__global__ void __launch_bounds__(64) k_energy(int n, int nq, int nchunk,
    const double *__restrict__ qpos, const double *__restrict__ src8, double eps, double kappa,
    double *__restrict__ part)
{
    const int task = blockIdx.x;
    const int g = task / nchunk, ch = task % nchunk;
    const int lane = lane_id();
    const int k = g * kWave + lane;
    const bool valid = k < nq;
    const double cx = valid ? qpos[3 * k] : 0.0, cy = valid ? qpos[3 * k + 1] : 0.0,
                 cz = valid ? qpos[3 * k + 2] : 0.0;
    const int f0 = ch * kEnergyChunk, f1 = min(n, f0 + kEnergyChunk);
    double s = 0.0;
#pragma unroll 2
    for (int j = f0; j < f1; ++j) {
        const double4 ra = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8);
        const double4 rb = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8 + 4);
        const double dx = ra.x - cx, dy = ra.y - cy, dz = ra.z - cz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double inv = rsq_nr(r2);
        const double rs = r2 * inv;
        const double krs = kappa * rs;
        const double ek = exp(-krs);
        const double g0 = kOneOver4Pi * inv;
        const double ct = (ra.w * dx + rb.x * dy + rb.y * dz) * inv;
        const double g1 = ct * g0 * inv;
        const double h1 = g1 * fma(-eps, fma(krs, ek, ek), 1.0);   // G1 - eps G2
        const double h2 = g0 - ek * g0;                            // G0 - Gk
        s = fma(h1, rb.z, fma(h2, rb.w, s));
    }
    part[(size_t)task * kWave + lane] = s;
}

// Mixed precision: the energy's source records in FP32, two float4 per face:
// {x, y, z} relative to the first face of the face's chunk (the difference
// taken in FP64, then rounded) and the normal {nx, ny, nz}; the .w entries
// hold the weights w1 = A*phi, w2 = A*dphi/dn of the solution when src8 holds
// them (weights true), else 0 (the block form reads its weights apart).
__global__ void k_energy_rec32(int n, const double *__restrict__ src8, bool weights,
                               float4 *__restrict__ rec)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double *a = src8 + (size_t)j * 8;
    const double *o = src8 + (size_t)(j / kEnergyChunk * kEnergyChunk) * 8;
    rec[2 * (size_t)j] = make_float4((float)(a[0] - o[0]), (float)(a[1] - o[1]), (float)(a[2] - o[2]),
                                     weights ? (float)a[6] : 0.f);
    rec[2 * (size_t)j + 1] = make_float4((float)a[3], (float)a[4], (float)a[5],
                                         weights ? (float)a[7] : 0.f);
}

// The FP32 factors of one (charge, face) pair of the energy, d = face - charge,
// r = |d|: P = (n.d)(1 - eps e^{-kr}(1 + kr))/r^3 and Q = (1 - e^{-kr})/r, so
// that w1 P + w2 Q is 4 pi (w1 (G1 - eps G2) + w2 (G0 - Gk)) of k_energy.  The
// charge (qx, qy, qz) is relative to the chunk's first face, like the record.
__device__ __forceinline__ void energy_pq32(const float4 ra, const float4 rb, float qx, float qy,
                                            float qz, float kappa, float eps, float &P, float &Q)
{
    const float dx = ra.x - qx, dy = ra.y - qy, dz = ra.z - qz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float krs = kappa * (r2 * inv);
    const float ek = __expf(-krs);
    const float nd = fmaf(rb.x, dx, fmaf(rb.y, dy, rb.z * dz));
    const float a = fmaf(-eps, fmaf(krs, ek, ek), 1.f);
    P = nd * (inv * inv) * a * inv;
    Q = (1.f - ek) * inv;
}

// The potential at each charge as k_energy computes it, with the pair terms in
// FP32 (energy_pq32): the terms of 16 faces are summed in FP32, those sums in
// FP64.  Tasks and partials as k_energy's.
constexpr int kEnergySub = 16;
__global__ void __launch_bounds__(64) k_energy32(int n, int nq, int nchunk,
    const double *__restrict__ qpos, const double *__restrict__ src8,
    const float4 *__restrict__ rec, float eps, float kappa, double *__restrict__ part)
{
    const int task = blockIdx.x;
    const int g = task / nchunk, ch = task % nchunk;
    const int lane = lane_id();
    const int k = g * kWave + lane;
    const bool valid = k < nq;
    const int f0 = ch * kEnergyChunk, f1 = min(n, f0 + kEnergyChunk);
    const double *o = src8 + (size_t)f0 * 8;
    const float qx = valid ? (float)(qpos[3 * k] - o[0]) : 0.f;
    const float qy = valid ? (float)(qpos[3 * k + 1] - o[1]) : 0.f;
    const float qz = valid ? (float)(qpos[3 * k + 2] - o[2]) : 0.f;
    double sd = 0.0;
    for (int j0 = f0; j0 < f1; j0 += kEnergySub) {
        const int j1 = min(f1, j0 + kEnergySub);
        float s = 0.f;
#pragma unroll 4
        for (int j = j0; j < j1; ++j) {
            const float4 ra = rec[2 * (size_t)j], rb = rec[2 * (size_t)j + 1];
            float P, Q;
            energy_pq32(ra, rb, qx, qy, qz, kappa, eps, P, Q);
            s = fmaf(ra.w, P, fmaf(rb.w, Q, s));
        }
        sd += (double)s;
    }
    part[(size_t)task * kWave + lane] = sd * kOneOver4Pi;
}

__global__ void k_energy_reduce(int nq, int nchunk, const double *__restrict__ part,
                                double *__restrict__ ptl)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const int g = k / kWave, lane = k % kWave;
    double s = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) s += part[((size_t)g * nchunk + ch) * kWave + lane];
    ptl[k] = s;
}

// Coulomb energy sum_{i<j} q_i q_j / (eps_p r_ij) (src/main.f90:149-153):
// workgroup (bi, bj) takes the charges i of block bi against the charges j > i
// of chunk bj, both 256 wide, the j read wave-uniformly; chunks wholly at or
// below the block contribute nothing.  One partial per workgroup, at
// part[bj * gridDim.x + bi], summed in that order by the caller.
__global__ void __launch_bounds__(256) k_coulomb(int nq, const double *__restrict__ qpos,
    const double *__restrict__ q, double *__restrict__ part)
{
    __shared__ double red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j0 = blockIdx.y * 256, j1 = min(nq, j0 + 256);
    double s = 0.0;
    if (j1 - 1 > (int)blockIdx.x * 256) {
        const bool vi = i < nq;
        const double xi = vi ? qpos[3 * i] : 0.0, yi = vi ? qpos[3 * i + 1] : 0.0,
                     zi = vi ? qpos[3 * i + 2] : 0.0;
        for (int j = j0; j < j1; ++j) {
            const double dx = xi - qpos[3 * j], dy = yi - qpos[3 * j + 1], dz = zi - qpos[3 * j + 2];
            const double v = q[j] * rsq_nr(dx * dx + dy * dy + dz * dz);
            s += (vi && j > i) ? v : 0.0;
        }
        s *= vi ? q[i] : 0.0;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// BLAS-1 for GMRES (deterministic two-pass reductions)
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 512;

__global__ void __launch_bounds__(256) k_dot_part(int64_t n, const double *__restrict__ a,
    const double *__restrict__ b, double *__restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += a[i] * b[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_part(int np, const double *__restrict__ part,
    double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// blocks of vectors (several right-hand sides, or the Krylov vectors of several
// GMRES runs): column `col` of a block is the vector act[col] of a strided set,
// base + act[col]*stride.  The field kernels evaluate each (target, source)
// kernel value and each (target, cluster) Taylor expansion once and apply it to
// RC columns held in registers.  Per column the arithmetic (operation order,
// fused multiply-adds, compensation) is that of the single-vector kernels above.
// ---------------------------------------------------------------------------
// source weights wb[j][2*Rp]: {A_j x[j], A_j x[n+j]} of column r < R, zero for
// the padding columns R <= r < Rp
template <typename T>
__global__ void k_src_weightsb(int n, int R, int Rp, const double *__restrict__ tarea,
    const double *__restrict__ x, int64_t xs, const int32_t *__restrict__ act, T *__restrict__ wb)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (j >= n) return;
    T w1 = (T)0, w2 = (T)0;
    if (r < R) {
        const double *xr = x + (size_t)act[r] * xs;
        const double a = tarea[j];
        w1 = (T)(a * xr[j]);
        w2 = (T)(a * xr[n + j]);
    }
    wb[(size_t)j * 2 * Rp + 2 * r] = w1;
    wb[(size_t)j * 2 * Rp + 2 * r + 1] = w2;
}

template <int RC>
__global__ void __launch_bounds__(64) k_nearb(const NearTask *__restrict__ tasks, 
    const int32_t *__restrict__ near_tgt, int ntask, int n,
    int R, int Rp, const double *__restrict__ src8, const double *__restrict__ wb,
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double kappa, double eps, double2 *__restrict__ part)
{
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const int r0 = blockIdx.y * RC;
    const NearTask tk = tasks[task];
    const int lane = lane_id();
    const int i = near_tgt[(size_t)task * kWave + lane];   // this lane's target, -1: none
    const bool act = i >= 0;
    const bool valid = act;
    const double sx = valid ? tx[i] : 0.0, sy = valid ? ty[i] : 0.0, sz = valid ? tz[i] : 0.0;
    const double v0x = valid ? tnx[i] : 0.0, v0y = valid ? tny[i] : 0.0,
                 v0z = valid ? tnz[i] : 0.0;
    const double kappa2 = kappa * kappa, inv_eps = 1.0 / eps;
    double p1[RC], p2[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) { p1[k] = 0.0; p2[k] = 0.0; }
    if (act) {
        for (int j = tk.s0; j < tk.s1; ++j) {
            const double4 ra = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8);
            const double4 rb = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8 + 4);
            const bool use = (j != i);
            const double dx = ra.x - sx, dy = ra.y - sy, dz = ra.z - sz;
            double r2 = dx * dx + dy * dy + dz * dz;
            r2 = use ? r2 : 1.0;
            const PairL L = pair_l(dx, dy, dz, r2, ra.w, rb.x, rb.y, v0x, v0y, v0z, kappa,
                                   kappa2, eps, inv_eps);
            const double *w = wb + (size_t)j * 2 * Rp + 2 * r0;
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                const double w1 = use ? w[2 * k] : 0.0, w2 = use ? w[2 * k + 1] : 0.0;
                p1[k] = fma(L.l1, w1, fma(L.l2, w2, p1[k]));
                p2[k] = fma(L.l3, w1, fma(L.l4, w2, p2[k]));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RC; ++k)
        if (r0 + k < R) part[((size_t)task * R + r0 + k) * kWave + lane] = make_double2(p1[k], p2[k]);
}

template <int RC>
__global__ void __launch_bounds__(64) k_near32b(const NearTask *__restrict__ tasks, 
    const int32_t *__restrict__ near_tgt, int ntask,
    int n, int R, int Rp, const float *__restrict__ src32, const float *__restrict__ wb,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, float kappa, float eps, double2 *__restrict__ part)
{
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const int r0 = blockIdx.y * RC;
    const NearTask tk = tasks[task];
    const int lane = lane_id();
    const int i = near_tgt[(size_t)task * kWave + lane];   // this lane's target, -1: none
    const bool act = i >= 0;
    const double4 lg = ngeo[tk.leaf];
    Near32 nk;
    nk.sx = act ? (float)(tx[i] - lg.x) : 0.f;
    nk.sy = act ? (float)(ty[i] - lg.y) : 0.f;
    nk.sz = act ? (float)(tz[i] - lg.z) : 0.f;
    nk.v0x = act ? (float)tnx[i] : 0.f;
    nk.v0y = act ? (float)tny[i] : 0.f;
    nk.v0z = act ? (float)tnz[i] : 0.f;
    nk.kappa = kappa; nk.kappa2 = kappa * kappa; nk.eps = eps; nk.inv_eps = 1.f / eps;
    Kahan p1[RC], p2[RC];
    if (act) {
        for (int j0 = tk.s0; j0 < tk.s1; j0 += kNearBlk) {
            float b1[RC], b2[RC];
#pragma unroll
            for (int k = 0; k < RC; ++k) b1[k] = b2[k] = 0.f;
#pragma unroll
            for (int u = 0; u < kNearBlk; ++u) {
                const int j = j0 + u;
                if (j < tk.s1) {
                    const float4 ra = *reinterpret_cast<const float4 *>(src32 + (size_t)j * 8);
                    const float4 rb = *reinterpret_cast<const float4 *>(src32 + (size_t)j * 8 + 4);
                    float l1, l2, l3, l4;
                    nk.terms(ra, rb, j != i, l1, l2, l3, l4);
                    const float *w = wb + (size_t)j * 2 * Rp + 2 * r0;
#pragma unroll
                    for (int k = 0; k < RC; ++k) {
                        b1[k] = fmaf(l1, w[2 * k], fmaf(l2, w[2 * k + 1], b1[k]));
                        b2[k] = fmaf(l3, w[2 * k], fmaf(l4, w[2 * k + 1], b2[k]));
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                p1[k].add(b1[k]);
                p2[k].add(b2[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RC; ++k)
        if (r0 + k < R)
            part[((size_t)task * R + r0 + k) * kWave + lane] =
                make_double2((double)p1[k].s, (double)p2[k].s);
}

// moment vectors of column r: mom_w + r*wstride (wstride = nmom*8*NW)
template <int P, int RC>
__global__ void __launch_bounds__(64) k_farb(const FarTask *__restrict__ tasks, int ntask, int n,
    int R, const int32_t *__restrict__ far_node, const u64 *__restrict__ far_mask,
    const int32_t *__restrict__ node2mom, const double *__restrict__ mom_w, int64_t wstride,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double kappa, double2 *__restrict__ part)
{
    constexpr int L = MomShape<P>::L;
    constexpr int NJ = MomShape<P>::NJ;
    constexpr int NW = MomShape<P>::NW;
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const int r0 = blockIdx.y * RC, rc = min(RC, R - r0);
    const FarTask tk = tasks[task];
    const int lane = lane_id();
    const int i = tk.g * kWave + lane;
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    double acc1[RC], ax[RC], ay[RC], az[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) { acc1[k] = 0.0; ax[k] = 0.0; ay[k] = 0.0; az[k] = 0.0; }
    for (int it = tk.i0; it < tk.i1; ++it) {
        const int nd = far_node[it];
        const u64 m = far_mask[it];
        if (!((m >> lane) & 1ull)) continue;
        const double4 geo = ngeo[nd];
        const double *W0 = mom_w + (size_t)r0 * wstride + (size_t)node2mom[nd] * 8 * NW;
        const double dx = px - geo.x, dy = py - geo.y, dz = pz - geo.z;
        {
            double a[NJ];
            tabi::taylor<L, false, double>(dx, dy, dz, 0.0, a);
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                if (k >= rc) break;
                const double *W = W0 + (size_t)k * wstride;
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    acc1[k] = fma(a[q], W[q], acc1[k]);
                    ax[k] = fma(a[tabi::tshift(q, 0)], W[2 * NW + q], ax[k]);
                    ay[k] = fma(a[tabi::tshift(q, 1)], W[3 * NW + q], ay[k]);
                    az[k] = fma(a[tabi::tshift(q, 2)], W[4 * NW + q], az[k]);
                }
            }
        }
        {
            double a[NJ];
            tabi::taylor<L, true, double>(dx, dy, dz, kappa, a);
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                if (k >= rc) break;
                const double *W = W0 + (size_t)k * wstride;
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    acc1[k] = fma(a[q], W[NW + q], acc1[k]);
                    ax[k] = fma(a[tabi::tshift(q, 0)], W[5 * NW + q], ax[k]);
                    ay[k] = fma(a[tabi::tshift(q, 1)], W[6 * NW + q], ay[k]);
                    az[k] = fma(a[tabi::tshift(q, 2)], W[7 * NW + q], az[k]);
                }
            }
        }
    }
    const double nx = valid ? tnx[i] : 0.0, ny = valid ? tny[i] : 0.0, nz = valid ? tnz[i] : 0.0;
#pragma unroll
    for (int k = 0; k < RC; ++k) {
        if (k >= rc) break;
        double p2 = 0.0;
        if (valid) p2 = nx * ax[k] + ny * ay[k] + nz * az[k];
        part[((size_t)task * R + r0 + k) * kWave + lane] = make_double2(acc1[k], p2);
    }
}

template <int P, int RC>
__global__ void __launch_bounds__(64) k_far32b(const FarTask *__restrict__ tasks, int ntask, int n,
    int R, const int32_t *__restrict__ far_node, const u64 *__restrict__ far_mask,
    const int32_t *__restrict__ node2mom, const float *__restrict__ mom_w, int64_t wstride,
    const double4 *__restrict__ ngeo, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, float kappa, double2 *__restrict__ part)
{
    constexpr int L = MomShape<P>::L;
    constexpr int NJ = MomShape<P>::NJ;
    constexpr int NW = MomShape<P>::NW;
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const int r0 = blockIdx.y * RC, rc = min(RC, R - r0);
    const FarTask tk = tasks[task];
    const int lane = lane_id();
    const int i = tk.g * kWave + lane;
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    const float nx = valid ? (float)tnx[i] : 0.f, ny = valid ? (float)tny[i] : 0.f,
                nz = valid ? (float)tnz[i] : 0.f;
    Kahan s1[RC], s2[RC];
    float acc1[RC], ax[RC], ay[RC], az[RC];
    for (int it = tk.i0; it < tk.i1; ++it) {
        const int nd = far_node[it];
        const u64 m = far_mask[it];
        if (!((m >> lane) & 1ull)) continue;
        const double4 geo = ngeo[nd];
        const float *W0 = mom_w + (size_t)r0 * wstride + (size_t)node2mom[nd] * 8 * NW;
        const float dx = (float)(px - geo.x), dy = (float)(py - geo.y), dz = (float)(pz - geo.z);
#pragma unroll
        for (int k = 0; k < RC; ++k) { acc1[k] = 0.f; ax[k] = 0.f; ay[k] = 0.f; az[k] = 0.f; }
        {
            float a[NJ];
            tabi::taylor<L, false, float>(dx, dy, dz, 0.f, a);
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                if (k >= rc) break;
                const float *W = W0 + (size_t)k * wstride;
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    acc1[k] = fmaf(a[q], W[q], acc1[k]);
                    ax[k] = fmaf(a[tabi::tshift(q, 0)], W[2 * NW + q], ax[k]);
                    ay[k] = fmaf(a[tabi::tshift(q, 1)], W[3 * NW + q], ay[k]);
                    az[k] = fmaf(a[tabi::tshift(q, 2)], W[4 * NW + q], az[k]);
                }
            }
        }
        {
            float a[NJ];
            tabi::taylor<L, true, float>(dx, dy, dz, kappa, a);
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                if (k >= rc) break;
                const float *W = W0 + (size_t)k * wstride;
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    acc1[k] = fmaf(a[q], W[NW + q], acc1[k]);
                    ax[k] = fmaf(a[tabi::tshift(q, 0)], W[5 * NW + q], ax[k]);
                    ay[k] = fmaf(a[tabi::tshift(q, 1)], W[6 * NW + q], ay[k]);
                    az[k] = fmaf(a[tabi::tshift(q, 2)], W[7 * NW + q], az[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RC; ++k) {
            if (k >= rc) break;
            s1[k].add(acc1[k]);
            s2[k].add(nx * ax[k] + ny * ay[k] + nz * az[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < RC; ++k) {
        if (k >= rc) break;
        part[((size_t)task * R + r0 + k) * kWave + lane] =
            make_double2((double)s1[k].s, (double)s2[k].s);
    }
}

// y = pre*x - sum of the target's partials, for block column blockIdx.y
__global__ void k_reduceb(int n, int R, NearSlots ns, const int32_t *__restrict__ far_task_off,
    const double2 *__restrict__ part_near, const double2 *__restrict__ part_far,
    const double *__restrict__ x, int64_t xs, double *__restrict__ y, int64_t ys,
    const int32_t *__restrict__ act, double pre1, double pre2)
{
    // item records loaded by the wave's lanes together, as in k_reduce
    __shared__ u64 s_mask[4][kWave];
    __shared__ int s_ipos[4][kWave], s_toff[4][kWave], s_nsp[4][kWave];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const int g = i / kWave, lane = i % kWave, w = threadIdx.x / kWave;
    if (g * kWave >= n) return;                  // the whole wave is past the end
    const u64 below = (1ull << lane) - 1;
    double p1 = 0.0, p2 = 0.0;
    const int k0 = ns.off[g], k1 = ns.off[g + 1];
    for (int kb = k0; kb < k1; kb += kWave) {
        const int cnt = min(kWave, k1 - kb);
        if (lane < cnt) {
            const int k = kb + lane, nd = ns.leaf[k];
            s_mask[w][lane] = ns.mask[k];
            s_ipos[w][lane] = ns.ipos[k];
            s_toff[w][lane] = ns.toff[nd];
            s_nsp[w][lane] = ns.nsplit[nd];
        }
        __builtin_amdgcn_wave_barrier();
        for (int kk = 0; kk < cnt; ++kk) {
            const u64 m = s_mask[w][kk];
            if (!((m >> lane) & 1)) continue;
            const int p = s_ipos[w][kk] + __popcll(m & below);
            const int nsp = s_nsp[w][kk];
            const size_t t0 = (size_t)s_toff[w][kk] + (size_t)(p / kWave) * nsp;
            for (int s = 0; s < nsp; ++s) {
                const double2 v = part_near[((t0 + s) * R + r) * kWave + p % kWave];
                p1 += v.x; p2 += v.y;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (i >= n) return;
    for (int t = far_task_off[g]; t < far_task_off[g + 1]; ++t) {
        const double2 v = part_far[((size_t)t * R + r) * kWave + lane];
        p1 += v.x; p2 += v.y;
    }
    const double *xr = x + (size_t)act[r] * xs;
    double *yr = y + (size_t)act[r] * ys;
    yr[i] = pre1 * xr[i] - p1;
    yr[n + i] = pre2 * xr[n + i] - p2;
}

// z = M^-1 r for RC block columns per workgroup (blockIdx.y): the factor
// stream is read once and applied to every column; z lives in LDS [RC][m]
template <int RC, typename FT>
__device__ __forceinline__ void gemv_rowb(const FT *__restrict__ Ac, int m, int bs,
                                          const double *zt, int ldz, double (&s)[RC])
{
    double s1[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) s1[k] = 0.0;
#pragma unroll 1
    for (int h = 0; h < kSolveB; h += kHalf) {
        FT af[kHalf];
#pragma unroll
        for (int c = 0; c < kHalf; ++c)
            af[c] = factor_ld(Ac, (size_t)(h + c) * m, (size_t)min(h + c, bs - 1) * m, h + c < bs);
        double a[kHalf];
#pragma unroll
        for (int c = 0; c < kHalf; ++c) a[c] = (double)af[c];
#pragma unroll
        for (int k = 0; k < RC; ++k) {
#pragma unroll
            for (int c = 0; c < kHalf; c += 2) {
                s[k] = fma(-a[c], (h + c < bs) ? zt[k * ldz + h + c] : 0.0, s[k]);
                s1[k] = fma(-a[c + 1], (h + c + 1 < bs) ? zt[k * ldz + h + c + 1] : 0.0, s1[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RC; ++k) s[k] = s[k] + s1[k];
}

template <int RC, typename FT>
__global__ void __launch_bounds__(64) k_lu_solveb(int n, int R, const int32_t *__restrict__ blk_ibeg,
    const int32_t *__restrict__ blk_m, const int64_t *__restrict__ blk_off,
    const int32_t *__restrict__ piv_off, const FT *__restrict__ lu,
    const int32_t *__restrict__ perm, const double *__restrict__ rv, int64_t rs_,
    double *__restrict__ zv, int64_t zs_, const int32_t *__restrict__ act, int ldz)
{
    extern __shared__ double zsh[];
    const int l = blockIdx.x, lane = threadIdx.x;
    const int r0 = blockIdx.y * RC, rc = min(RC, R - r0);
    const int m = blk_m[l], nl = m / 2, ib = blk_ibeg[l];
    const FT *A = lu + blk_off[l];
    const int32_t *pm = perm + piv_off[l];
    for (int k = 0; k < RC; ++k) {
        const double *r = k < rc ? rv + (size_t)act[r0 + k] * rs_ : nullptr;
        for (int q = lane; q < m; q += kWave) {
            const int src = pm[q];
            zsh[k * ldz + q] = k < rc ? (src < nl ? r[ib + src] : r[n + ib + src - nl]) : 0.0;
        }
    }
    __syncthreads();
    const int nb = (m + kSolveB - 1) / kSolveB;
    for (int bi = 0; bi < nb; ++bi) {
        const int kb = bi * kSolveB, bs = min(kSolveB, m - kb);
        const bool in = lane < bs;
        double zl[RC];
#pragma unroll
        for (int k = 0; k < RC; ++k) zl[k] = in ? zsh[k * ldz + kb + lane] : 0.0;
#pragma unroll 1
        for (int h = 0; h < kSolveB; h += kHalf) {
            FT tf[kHalf];
#pragma unroll
            for (int c = 0; c < kHalf; ++c)
                tf[c] = factor_ld(A, (size_t)(kb + h + c) * m + kb + lane,
                                  (size_t)(kb + min(h + c, bs - 1)) * m + kb + min(lane, bs - 1),
                                  in && h + c < lane);
            double tr[kHalf];
#pragma unroll
            for (int c = 0; c < kHalf; ++c) tr[c] = (double)tf[c];
#pragma unroll
            for (int k = 0; k < RC; ++k)
#pragma unroll
                for (int c = 0; c < kHalf; ++c) zl[k] -= tr[c] * readlane_d(zl[k], h + c);
        }
        if (in) {
#pragma unroll
            for (int k = 0; k < RC; ++k) zsh[k * ldz + kb + lane] = zl[k];
        }
        __syncthreads();
        for (int q0 = kb + bs; q0 < m; q0 += kWave) {
            const int row = q0 + lane;
            if (row < m) {
                double sv[RC];
#pragma unroll
                for (int k = 0; k < RC; ++k) sv[k] = zsh[k * ldz + row];
                gemv_rowb<RC, FT>(A + (size_t)kb * m + row, m, bs, zsh + kb, ldz, sv);
#pragma unroll
                for (int k = 0; k < RC; ++k) zsh[k * ldz + row] = sv[k];
            }
        }
        __syncthreads();
    }
    for (int bi = nb - 1; bi >= 0; --bi) {
        const int kb = bi * kSolveB, bs = min(kSolveB, m - kb);
        const bool in = lane < bs;
        double zl[RC];
#pragma unroll
        for (int k = 0; k < RC; ++k) zl[k] = in ? zsh[k * ldz + kb + lane] : 0.0;
#pragma unroll 1
        for (int h = kSolveB - kHalf; h >= 0; h -= kHalf) {
            FT tf[kHalf];
#pragma unroll
            for (int c = 0; c < kHalf; ++c)
                tf[c] = factor_ld(A, (size_t)(kb + h + c) * m + kb + lane,
                                  (size_t)(kb + min(h + c, bs - 1)) * m + kb + min(lane, bs - 1),
                                  in && h + c >= lane && h + c < bs);
            double tr[kHalf];
#pragma unroll
            for (int c = 0; c < kHalf; ++c) tr[c] = (double)tf[c];
#pragma unroll
            for (int c = kHalf - 1; c >= 0; --c) {
#pragma unroll
                for (int k = 0; k < RC; ++k) {
                    if (lane == h + c && in) zl[k] = zl[k] / tr[c];
                    const double zc = readlane_d(zl[k], h + c);
                    if (lane < h + c) zl[k] -= tr[c] * zc;
                }
            }
        }
        if (in) {
#pragma unroll
            for (int k = 0; k < RC; ++k) zsh[k * ldz + kb + lane] = zl[k];
        }
        __syncthreads();
        for (int q0 = 0; q0 < kb; q0 += kWave) {
            double sv[RC];
#pragma unroll
            for (int k = 0; k < RC; ++k) sv[k] = zsh[k * ldz + q0 + lane];
            gemv_rowb<RC, FT>(A + (size_t)kb * m + q0 + lane, m, bs, zsh + kb, ldz, sv);
#pragma unroll
            for (int k = 0; k < RC; ++k) zsh[k * ldz + q0 + lane] = sv[k];
        }
        __syncthreads();
    }
    for (int k = 0; k < rc; ++k) {
        double *z = zv + (size_t)act[r0 + k] * zs_;
        for (int q = lane; q < nl; q += kWave) {
            z[ib + q] = zsh[k * ldz + q];
            z[n + ib + q] = zsh[k * ldz + nl + q];
        }
    }
}

// right-hand sides of R charge sets at the same positions (charges q[r][k]):
// the geometric factors are formed once per (face, charge) and applied to RC
// sets per workgroup column (blockIdx.y)
template <int RC>
__global__ void __launch_bounds__(256) k_rhsb(int n, int nq, int R, const double *__restrict__ qpos,
    const double *__restrict__ q, double eps_p, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double *__restrict__ b)
{
    __shared__ double sq[3 + RC][256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * RC, rc = min(RC, R - r0);
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    const double vx = valid ? tnx[i] : 0.0, vy = valid ? tny[i] : 0.0, vz = valid ? tnz[i] : 0.0;
    double b1[RC], b2[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) { b1[k] = 0.0; b2[k] = 0.0; }
    for (int k0 = 0; k0 < nq; k0 += 256) {
        const int kk = k0 + threadIdx.x;
        __syncthreads();
        if (kk < nq) {
            sq[0][threadIdx.x] = qpos[3 * kk];
            sq[1][threadIdx.x] = qpos[3 * kk + 1];
            sq[2][threadIdx.x] = qpos[3 * kk + 2];
#pragma unroll
            for (int k = 0; k < RC; ++k) sq[3 + k][threadIdx.x] = k < rc ? q[(size_t)(r0 + k) * nq + kk] : 0.0;
        }
        __syncthreads();
        const int cnt = min(256, nq - k0);
        for (int c = 0; c < cnt; ++c) {
            const double dx = sq[0][c] - px, dy = sq[1][c] - py, dz = sq[2][c] - pz;
            const double inv = rsq_nr(dx * dx + dy * dy + dz * dz);
            const double ct = (vx * dx + vy * dy + vz * dz) * inv;
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                b1[k] += sq[3 + k][c] * inv;
                b2[k] += sq[3 + k][c] * ct * inv * inv;
            }
        }
    }
    if (valid) {
        const double s = 1.0 / (4.0 * kPi * eps_p);
#pragma unroll
        for (int k = 0; k < RC; ++k) {
            if (k >= rc) break;
            double *br = b + (size_t)(r0 + k) * 2 * n;
            br[i] = b1[k] * s;
            br[n + i] = b2[k] * s;
        }
    }
}

// k_rhsb with FP32 pair terms (mixed precision), summed as k_rhs32 sums them,
// so each set's right-hand side equals its single-set one
template <int RC>
__global__ void __launch_bounds__(256) k_rhsb32(int n, int nq, int R, const double *__restrict__ qpos,
    const double *__restrict__ q, double eps_p, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ tnx, const double *__restrict__ tny,
    const double *__restrict__ tnz, double *__restrict__ b)
{
    static_assert(256 % kRhsSub == 0, "charge batches must hold whole FP32 partials");
    __shared__ double sq[3][256];
    __shared__ float sqq[RC][256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * RC, rc = min(RC, R - r0);
    const bool valid = i < n;
    const double px = valid ? tx[i] : 0.0, py = valid ? ty[i] : 0.0, pz = valid ? tz[i] : 0.0;
    const float vx = valid ? (float)tnx[i] : 0.f, vy = valid ? (float)tny[i] : 0.f,
                vz = valid ? (float)tnz[i] : 0.f;
    double b1[RC], b2[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) { b1[k] = 0.0; b2[k] = 0.0; }
    for (int k0 = 0; k0 < nq; k0 += 256) {
        const int kk = k0 + threadIdx.x;
        __syncthreads();
        if (kk < nq) {
            sq[0][threadIdx.x] = qpos[3 * kk];
            sq[1][threadIdx.x] = qpos[3 * kk + 1];
            sq[2][threadIdx.x] = qpos[3 * kk + 2];
#pragma unroll
            for (int k = 0; k < RC; ++k)
                sqq[k][threadIdx.x] = k < rc ? (float)q[(size_t)(r0 + k) * nq + kk] : 0.f;
        }
        __syncthreads();
        const int cnt = min(256, nq - k0);
        for (int c0 = 0; c0 < cnt; c0 += kRhsSub) {
            const int c1 = min(cnt, c0 + kRhsSub);
            float s1[RC], s2[RC];
#pragma unroll
            for (int k = 0; k < RC; ++k) { s1[k] = 0.f; s2[k] = 0.f; }
            for (int c = c0; c < c1; ++c) {
                const float dx = (float)(sq[0][c] - px), dy = (float)(sq[1][c] - py),
                            dz = (float)(sq[2][c] - pz);
                const float inv = __builtin_amdgcn_rsqf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
                const float g = fmaf(vx, dx, fmaf(vy, dy, vz * dz)) * (inv * inv);
#pragma unroll
                for (int k = 0; k < RC; ++k) {
                    const float qi = sqq[k][c] * inv;
                    s1[k] += qi;
                    s2[k] = fmaf(qi, g, s2[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < RC; ++k) { b1[k] += (double)s1[k]; b2[k] += (double)s2[k]; }
        }
    }
    if (valid) {
        const double s = 1.0 / (4.0 * kPi * eps_p);
#pragma unroll
        for (int k = 0; k < RC; ++k) {
            if (k >= rc) break;
            double *br = b + (size_t)(r0 + k) * 2 * n;
            br[i] = b1[k] * s;
            br[n + i] = b2[k] * s;
        }
    }
}

// potential at each charge from R surface solutions (weights wb[j][2*Rp] as
// k_src_weightsb writes them); tasks as k_energy's, RC columns per blockIdx.y
template <int RC>
__global__ void __launch_bounds__(64) k_energyb(int n, int nq, int nchunk, int R, int Rp,
    const double *__restrict__ qpos, const double *__restrict__ src8,
    const double *__restrict__ wb, double eps, double kappa, double *__restrict__ part)
{
    const int task = blockIdx.x;
    const int r0 = blockIdx.y * RC;
    const int g = task / nchunk, ch = task % nchunk;
    const int lane = lane_id();
    const int k = g * kWave + lane;
    const bool valid = k < nq;
    const double cx = valid ? qpos[3 * k] : 0.0, cy = valid ? qpos[3 * k + 1] : 0.0,
                 cz = valid ? qpos[3 * k + 2] : 0.0;
    const int f0 = ch * kEnergyChunk, f1 = min(n, f0 + kEnergyChunk);
    double s[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) s[c] = 0.0;
    for (int j = f0; j < f1; ++j) {
        const double4 ra = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8);
        const double4 rb = *reinterpret_cast<const double4 *>(src8 + (size_t)j * 8 + 4);
        const double dx = ra.x - cx, dy = ra.y - cy, dz = ra.z - cz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double inv = rsq_nr(r2);
        const double rs = r2 * inv;
        const double krs = kappa * rs;
        const double ek = exp(-krs);
        const double g0 = kOneOver4Pi * inv;
        const double ct = (ra.w * dx + rb.x * dy + rb.y * dz) * inv;
        const double g1 = ct * g0 * inv;
        const double h1 = g1 * fma(-eps, fma(krs, ek, ek), 1.0);
        const double h2 = g0 - ek * g0;
        const double *w = wb + (size_t)j * 2 * Rp + 2 * r0;
#pragma unroll
        for (int c = 0; c < RC; ++c) s[c] = fma(h1, w[2 * c], fma(h2, w[2 * c + 1], s[c]));
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
        if (r0 + c < R) part[((size_t)task * R + r0 + c) * kWave + lane] = s[c];
}

// k_energyb with the pair terms in FP32 (energy_pq32, summed as k_energy32
// sums them, so a column's potentials equal those of a single solve's);
// records from k_energy_rec32 without weights, weights wbf[j][2*Rp] in FP32.
template <int RC>
__global__ void __launch_bounds__(64) k_energyb32(int n, int nq, int nchunk, int R, int Rp,
    const double *__restrict__ qpos, const double *__restrict__ src8,
    const float4 *__restrict__ rec, const float *__restrict__ wb, float eps, float kappa,
    double *__restrict__ part)
{
    const int task = blockIdx.x;
    const int r0 = blockIdx.y * RC;
    const int g = task / nchunk, ch = task % nchunk;
    const int lane = lane_id();
    const int k = g * kWave + lane;
    const bool valid = k < nq;
    const int f0 = ch * kEnergyChunk, f1 = min(n, f0 + kEnergyChunk);
    const double *o = src8 + (size_t)f0 * 8;
    const float qx = valid ? (float)(qpos[3 * k] - o[0]) : 0.f;
    const float qy = valid ? (float)(qpos[3 * k + 1] - o[1]) : 0.f;
    const float qz = valid ? (float)(qpos[3 * k + 2] - o[2]) : 0.f;
    double sd[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) sd[c] = 0.0;
    for (int j0 = f0; j0 < f1; j0 += kEnergySub) {
        const int j1 = min(f1, j0 + kEnergySub);
        float s[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) s[c] = 0.f;
#pragma unroll 2
        for (int j = j0; j < j1; ++j) {
            const float4 ra = rec[2 * (size_t)j], rb = rec[2 * (size_t)j + 1];
            float P, Q;
            energy_pq32(ra, rb, qx, qy, qz, kappa, eps, P, Q);
            const float *w = wb + (size_t)j * 2 * Rp + 2 * r0;
#pragma unroll
            for (int c = 0; c < RC; ++c) s[c] = fmaf(w[2 * c], P, fmaf(w[2 * c + 1], Q, s[c]));
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) sd[c] += (double)s[c];
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
        if (r0 + c < R) part[((size_t)task * R + r0 + c) * kWave + lane] = sd[c] * kOneOver4Pi;
}

__global__ void k_energy_reduceb(int nq, int nchunk, int R, const double *__restrict__ part,
                                 double *__restrict__ ptl)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (k >= nq) return;
    const int g = k / kWave, lane = k % kWave;
    double s = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) s += part[(((size_t)g * nchunk + ch) * R + r) * kWave + lane];
    ptl[(size_t)r * nq + k] = s;
}

// Batched BLAS-1 for R GMRES runs advanced together.  Run r's vectors are
// found through a pointer table, tab[slot*R + r], and have length nv[r]; the
// column list act maps grid column blockIdx.y to a run.  Per run the arithmetic
// is that of the single-vector kernels (same partition and summation order).
// A vector may lie in two segments: its second half (the normal-derivative
// unknowns) starts nv[R + r] elements after the end of its first half, which
// is how a run's vector sits inside a vector of several packed surfaces
// ([all potentials; all normal derivatives]); nv[R + r] = 0 for a contiguous
// vector.
__device__ __forceinline__ int64_t vix(int64_t i, int64_t half, int64_t gap)
{
    return i + (i >= half ? gap : 0);
}
// y += (sgn * coef[r*cs + cofs]) * x
__global__ void k_axpyt(double *const *__restrict__ tab, int R, int sx, int sy,
    const int64_t *__restrict__ nv, const double *__restrict__ coef, int64_t cs, int cofs,
    double sgn, const int32_t *__restrict__ act, const int32_t *__restrict__ nact = nullptr)
{
    if (nact && (int)blockIdx.y >= *nact) return;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = act[blockIdx.y];
    if (i0 >= nv[r]) return;
    const int64_t i = vix(i0, nv[r] >> 1, nv[R + r]);
    const double a = sgn * coef[(size_t)r * cs + cofs];
    double *y = tab[(size_t)sy * R + r];
    y[i] += a * tab[(size_t)sx * R + r][i];
}

// x *= coef[r*cs + cofs]
__global__ void k_scalt(double *const *__restrict__ tab, int R, int sx, const int64_t *__restrict__ nv,
    const double *__restrict__ coef, int64_t cs, int cofs, const int32_t *__restrict__ act,
    const int32_t *__restrict__ nact = nullptr)
{
    if (nact && (int)blockIdx.y >= *nact) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = act[blockIdx.y];
    if (i < nv[r]) tab[(size_t)sx * R + r][vix(i, nv[r] >> 1, nv[R + r])] *= coef[(size_t)r * cs + cofs];
}

// out = sum_{k < nvv} coef[r*cs + k] * slot(sv0 + k)
// (with nvr, run r combines nvr[r] vectors instead of nvv)
__global__ void k_lincombt(double *const *__restrict__ tab, int R, int sv0, int nvv,
    const int64_t *__restrict__ nv, const double *__restrict__ coef, int64_t cs, int so,
    const int32_t *__restrict__ act, const int32_t *__restrict__ nact = nullptr,
    const int32_t *__restrict__ nvr = nullptr)
{
    if (nact && (int)blockIdx.y >= *nact) return;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = act[blockIdx.y];
    if (nvr) nvv = nvr[r];
    if (i0 >= nv[r]) return;
    const int64_t i = vix(i0, nv[r] >> 1, nv[R + r]);
    const double *cr = coef + (size_t)r * cs;
    double s = 0.0;
    for (int k = 0; k < nvv; ++k) s += cr[k] * tab[(size_t)(sv0 + k) * R + r][i];
    tab[(size_t)so * R + r][i] = s;
}

// Fused forms for the GMRES inner loop, with the per-element arithmetic, the
// partition (k_dot_part's grid-stride over gridDim.x = kRedBlocks blocks) and
// the summation order (k_sum_part's) of the unfused sequence (axpy, dot
// partials, sum of partials), so their results are bit-identical to it.  A dot
// leaves its kRedBlocks partials in part[(r*dvs + slot)*kRedBlocks ...]; a
// kernel that needs the value (the next step's axpy coefficient) sums them
// again in every block, in the same order, so no block waits for another and
// no cross-XCD fence is needed; k_sum_slots forms the values the host reads.

// elements per load group of a thread: vectors of up to kFuseU * kRedBlocks *
// 256 (~1M) entries take a single group
constexpr int kFuseU = 8;

// sum of np partials in k_sum_part's order (all 256 threads of the block)
__device__ __forceinline__ double sum_partials(const double *p, int np, double *red)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += p[i];
    __syncthreads();
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// y = slot sy.  When sx >= 0, first y -= h * slot(sx) (k_axpyt with sgn -1)
// where h is the dot of slot hs, summed from its partials.  Then the partials of
// slot(sa) . y into slot o1 when o1 >= 0, and of y . y into slot o2 when
// o2 >= 0.  Grid (kRedBlocks, columns).
// With hcond (the re-orthogonalisation pass of DORTH, src/dgmres_dep.f:1386-1411),
// the axpy is skipped when h + 0.001*c == h for h = hcond[r * hstride]: c is
// then DORTH's TEM, and h the Hessenberg entry it would change.
__device__ __forceinline__ bool reo_negligible(double h, double tem)
{
#pragma clang fp contract(off)
    return h + 0.001 * tem == h;
}

__global__ void __launch_bounds__(256) k_axpy_dott(double *const *__restrict__ tab, int R, int sx,
    int sy, int sa, const int64_t *__restrict__ nv, int hs, const int32_t *__restrict__ act,
    double *__restrict__ part, int64_t dvs, int o1, int o2, const int32_t *__restrict__ nact = nullptr,
    const double *__restrict__ hcond = nullptr, int64_t hstride = 0)
{
    __shared__ double red[2][256];
    if (nact && (int)blockIdx.y >= *nact) return;
    const int col = blockIdx.y, r = act[col], t = threadIdx.x;
    double *y = tab[(size_t)sy * R + r];
    const double *x = tab[(size_t)(sx >= 0 ? sx : sy) * R + r];
    const double *a = tab[(size_t)(o1 >= 0 ? sa : sy) * R + r];
    const int64_t n = nv[r], hn = n >> 1, gp = nv[R + r];
    double *pr = part + (size_t)r * dvs * gridDim.x;
    double c = sx >= 0 ? -1.0 * sum_partials(pr + (size_t)hs * gridDim.x, gridDim.x, red[0]) : 0.0;
    if (hcond && sx >= 0 && reo_negligible(hcond[(size_t)r * hstride], c)) sx = -1;
    double s1 = 0.0, s2 = 0.0;
    // a thread's elements i0 + u*stride in order, kFuseU at a time: all loads
    // of a group are issued before its stores, so they are in flight together
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i0 = (int64_t)blockIdx.x * 256 + t; i0 < n; i0 += stride * kFuseU) {
        double yv[kFuseU], xv[kFuseU], av[kFuseU];
#pragma unroll
        for (int u = 0; u < kFuseU; ++u) {
            const int64_t i = i0 + u * stride, j = vix(i, hn, gp);
            const bool ok = i < n;
            yv[u] = ok ? y[j] : 0.0;
            xv[u] = ok && sx >= 0 ? x[j] : 0.0;
            av[u] = ok && o1 >= 0 && sa != sy ? a[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kFuseU; ++u) {
            const int64_t i = i0 + u * stride;
            if (i >= n) break;
            double yi = yv[u];
            if (sx >= 0) {
                yi += c * xv[u];
                y[vix(i, hn, gp)] = yi;
            }
            if (o1 >= 0) s1 += (sa == sy ? yi : av[u]) * yi;
            if (o2 >= 0) s2 += yi * yi;
        }
    }
    __syncthreads();
    red[0][t] = s1;
    red[1][t] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) { red[0][t] += red[0][t + w]; red[1][t] += red[1][t + w]; }
        __syncthreads();
    }
    if (t == 0) {
        if (o1 >= 0) pr[(size_t)o1 * gridDim.x + blockIdx.x] = red[0][0];
        if (o2 >= 0) pr[(size_t)o2 * gridDim.x + blockIdx.x] = red[1][0];
    }
}

// Coefficients of one run passed by value (kernel arguments), for launches
// over a single run: no upload precedes them.
constexpr int kCoefArg = 32;
struct CoefArg {
    double v[kCoefArg];
};

// out = sum_{k < nvv} coef[r*cs + k] * slot(sv0 + k) (k_lincombt), the
// partials of out . out into slot o, and slot(sv0 + nvv - 1) *= coef[r*cs +
// csc] when csc >= 0 (k_scalt, after the slot was read for the combination).
// With by_value the single run's coefficients come from ca instead.
__global__ void __launch_bounds__(256) k_lincomb_dott(double *const *__restrict__ tab, int R, int sv0,
    int nvv, const int64_t *__restrict__ nv, const double *__restrict__ coef, int64_t cs, int so,
    int csc, const int32_t *__restrict__ act, double *__restrict__ part, int64_t dvs, int o,
    bool by_value, CoefArg ca, const int32_t *__restrict__ nact = nullptr)
{
    __shared__ double red[256];
    if (nact && (int)blockIdx.y >= *nact) return;
    const int col = blockIdx.y, r = act[col], t = threadIdx.x;
    const double *cr = by_value ? ca.v : coef + (size_t)r * cs;
    double *out = tab[(size_t)so * R + r];
    double *vl = tab[(size_t)(sv0 + nvv - 1) * R + r];
    const double sc = csc >= 0 ? cr[csc] : 1.0;
    const int64_t n = nv[r], hn = n >> 1, gp = nv[R + r];
    double s2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i0 = (int64_t)blockIdx.x * 256 + t; i0 < n; i0 += stride * kFuseU) {
        double sv[kFuseU], lv[kFuseU];
#pragma unroll
        for (int u = 0; u < kFuseU; ++u) sv[u] = 0.0;
        for (int k = 0; k < nvv; ++k) {
            const double ck = cr[k], *vk = tab[(size_t)(sv0 + k) * R + r];
#pragma unroll
            for (int u = 0; u < kFuseU; ++u) {
                const int64_t i = i0 + u * stride;
                const double v = i < n ? vk[vix(i, hn, gp)] : 0.0;
                sv[u] += ck * v;
                if (k == nvv - 1) lv[u] = v;
            }
        }
#pragma unroll
        for (int u = 0; u < kFuseU; ++u) {
            const int64_t i = i0 + u * stride;
            if (i >= n) break;
            const int64_t j = vix(i, hn, gp);
            out[j] = sv[u];
            if (csc >= 0) vl[j] = lv[u] * sc;
            s2 += sv[u] * sv[u];
        }
    }
    red[t] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) part[((size_t)r * dvs + o) * gridDim.x + blockIdx.x] = red[0];
}

// hval[r*dvs + slot] = sum of the slot's np partials (k_sum_part's order) for
// the slots [s0, s1) and the extra slots e0, e1 (when >= 0); one block per
// column.  hval and hflag are host memory (pinned, coherent, mapped): after
// its values, each block publishes hflag[r] = seq with system-scope release,
// and the host polls the flag instead of waiting for a copy.
__global__ void __launch_bounds__(256) k_sum_slots(int np, const double *__restrict__ part,
    double *__restrict__ hval, int64_t *__restrict__ hflag, int64_t seq, int64_t dvs, int s0, int s1,
    int e0, int e1, const int32_t *__restrict__ act)
{
    __shared__ double red[256];
    const int r = act[blockIdx.x];
    const double *pr = part + (size_t)r * dvs * np;
    for (int k = s0; k < s1 + 2; ++k) {
        const int slot = k < s1 ? k : k == s1 ? e0 : e1;
        if (slot < 0) continue;
        const double v = sum_partials(pr + (size_t)slot * np, np, red);
        if (threadIdx.x == 0) hval[(size_t)r * dvs + slot] = v;
    }
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(hflag + r, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}


// out = a - b
__global__ void k_subt(double *const *__restrict__ tab, int R, int sa, int sb, int so,
    const int64_t *__restrict__ nv, const int32_t *__restrict__ act)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = act[blockIdx.y];
    if (i >= nv[r]) return;
    const int64_t j = vix(i, nv[r] >> 1, nv[R + r]);
    tab[(size_t)so * R + r][j] = tab[(size_t)sa * R + r][j] - tab[(size_t)sb * R + r][j];
}

__global__ void k_copyt(double *const *__restrict__ tab, int R, int sa, int so,
    const int64_t *__restrict__ nv, const int32_t *__restrict__ act,
    const int32_t *__restrict__ nact = nullptr)
{
    if (nact && (int)blockIdx.y >= *nact) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = act[blockIdx.y];
    if (i < nv[r]) {
        const int64_t j = vix(i, nv[r] >> 1, nv[R + r]);
        tab[(size_t)so * R + r][j] = tab[(size_t)sa * R + r][j];
    }
}

__global__ void k_fillt(double *const *__restrict__ tab, int R, int so, const int64_t *__restrict__ nv,
    double v, const int32_t *__restrict__ act, const int32_t *__restrict__ nact = nullptr)
{
    if (nact && (int)blockIdx.y >= *nact) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = act[blockIdx.y];
    if (i < nv[r]) tab[(size_t)so * R + r][vix(i, nv[r] >> 1, nv[R + r])] = v;
}


// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
template <int P>
void launch_moments(tabi_ctx *c, const double *x)
{
    if (c->nmom == 0) return;
    k_mom_raw<P><<<c->nchunk, 256, 0, c->s>>>(c->n, c->chunk_node.p, c->chunk_beg.p,
        c->chunk_end.p, c->ngeo.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p,
        c->tarea.p, x, c->mom_part.p, c->gate, c->chunk_sf.p);
    k_mom_final<P><<<c->nmom, 256, 0, c->s>>>(c->mom_chunk_off.p, c->mom_part.p, c->prm.eps,
        c->mom_w.p, c->prm.precision == 1 ? c->mom_w32.p : nullptr, c->gate, c->mom_sf.p);
}

template <int P>
void launch_far(tabi_ctx *c)
{
    if (c->n_far_task == 0) return;
    if (c->prm.precision == 1) {
        // mom_w32 was written beside mom_w by k_mom_final
        k_far32<P><<<c->n_far_task, 64, 0, c->s>>>(c->far_tasks.p, c->n_far_task, c->n,
            c->far_node.p, c->far_mask.p, c->node2mom.p, c->mom_w32.p, c->ngeo.p, c->tx.p,
            c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, (float)c->prm.kappa, c->part_far.p,
            c->gate);
        return;
    }
    k_far<P><<<c->n_far_task, 64, 0, c->s>>>(c->far_tasks.p, c->n_far_task, c->n,
        c->far_node.p, c->far_mask.p, c->node2mom.p, c->mom_w.p, c->ngeo.p, c->tx.p, c->ty.p,
        c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, c->prm.kappa, c->part_far.p, c->gate);
}

void dispatch_moments(tabi_ctx *c, const double *x)
{
    switch (c->p) {
    case 1: launch_moments<1>(c, x); break;
    case 2: launch_moments<2>(c, x); break;
    case 3: launch_moments<3>(c, x); break;
    case 4: launch_moments<4>(c, x); break;
    case 5: launch_moments<5>(c, x); break;
    case 6: launch_moments<6>(c, x); break;
    default: throw Error("unsupported Taylor order");
    }
}

void dispatch_far(tabi_ctx *c)
{
    switch (c->p) {
    case 1: launch_far<1>(c); break;
    case 2: launch_far<2>(c); break;
    case 3: launch_far<3>(c); break;
    case 4: launch_far<4>(c); break;
    case 5: launch_far<5>(c); break;
    case 6: launch_far<6>(c); break;
    default: throw Error("unsupported Taylor order");
    }
}

// y = A x on device vectors c->vx -> c->vy
NearSlots near_slots(const tabi_ctx *c)
{
    return {c->near_off.p, c->near_leaf.p, c->near_mask.p, c->near_ipos.p, c->node_toff.p,
            c->node_nsplit.p};
}

void matvec_dev(tabi_ctx *c, const double *dx, double *dy)
{
    // the kernels read x in place; an in-place product (dy == dx) reads a copy
    const double *x = dx;
    if (dx == dy) {
        HIPCHK(hipMemcpyAsync(c->vx.p, dx, sizeof(double) * 2 * c->n, hipMemcpyDeviceToDevice, c->s));
        x = c->vx.p;
    }
    // The near field reads x and the geometry only, so it can run beside the
    // moments (short, latency-bound launches) and the far field on a second
    // stream, which fills the CUs those leave idle; the two write separate
    // partial arrays, so the product is the same either way.  Under graph
    // capture the fork and the join become the graph's edges.  Serial on one
    // stream when kernels are timed one by one.
    const bool side = c->n_near_task > 0 && c->nmom > 0 && c->n_far_task > 0 && !c->prof;
    hipStream_t ns = c->s;
    if (side) {
        c->aux.ensure();
        ns = c->aux.s;
        HIPCHK(hipEventRecord(c->aux.fork, c->s));
        HIPCHK(hipStreamWaitEvent(ns, c->aux.fork, 0));
    } else {
        KTimer kt(c, 0);
        dispatch_moments(c, x);
    }
    if (c->n_near_task) {
        KTimer kt(c, 1);
        if (c->prm.precision == 1) {
            k_src32_weights<<<nblk(c->n, 256), 256, 0, ns>>>(c->n, c->tarea.p, x, c->src32.p, c->gate,
                                                             c->group_sf.p);
            k_near32<<<c->n_near_task, 64, 0, ns>>>(c->near_tasks.p, c->near_tgt.p, c->n_near_task, c->n,
                c->src32.p, c->ngeo.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p,
                (float)c->prm.kappa, (float)c->prm.eps, c->part_near.p, c->gate);
        } else {
            k_src_weights<<<nblk(c->n, 256), 256, 0, ns>>>(c->n, c->tarea.p, x, c->src8.p, c->gate,
                                                           c->group_sf.p);
            k_near<<<c->n_near_task, 64, 0, ns>>>(c->near_tasks.p, c->near_tgt.p, c->n_near_task, c->n,
                c->src8.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, c->prm.kappa,
                c->prm.eps, c->part_near.p, c->gate);
        }
    }
    if (side) {
        HIPCHK(hipEventRecord(c->aux.join, ns));
        dispatch_moments(c, x);
    }
    if (c->n_far_task) { KTimer kt(c, 2); dispatch_far(c); }
    if (side) HIPCHK(hipStreamWaitEvent(c->s, c->aux.join, 0));
    const double pre1 = 0.5 * (1.0 + c->prm.eps), pre2 = 0.5 * (1.0 + 1.0 / c->prm.eps);
    {
        KTimer kt(c, 3);
        k_reduce<<<nblk(c->n, 256), 256, 0, c->s>>>(c->n, near_slots(c), c->far_task_off.p,
            c->part_near.p, c->part_far.p, x, pre1, pre2, dy, c->gate, c->group_sf.p);
    }
    HIPCHK(hipGetLastError());
}

// Diagonal scaling, MSOLVEprec (src/main.f90:532-542): z = r / (1/2)(1+eps) on
// the potential half and r / (1/2)(1+1/eps) on the normal-derivative half,
// divided as the reference divides (bit-identical).  Vector y of a block is
// act[y] (all vectors in order without a list), stored at stride rs / zs.
__global__ void k_diag_prec(int n, double eps, const double *__restrict__ rv, int64_t rs,
                            double *__restrict__ zv, int64_t zs, const int32_t *__restrict__ act)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * (int64_t)n) return;
    const int64_t k = act ? act[blockIdx.y] : blockIdx.y;
    const double scale1 = 0.5 * (1.0 + eps), scale2 = 0.5 * (1.0 + 1.0 / eps);
    zv[k * zs + i] = rv[k * rs + i] / (i < n ? scale1 : scale2);
}

void msolve_dev(tabi_ctx *c, const double *dr, double *dz)
{
    if (c->prm.flags & 4) {
        KTimer kt(c, 4);
        k_diag_prec<<<nblk(2 * (int64_t)c->n, 256), 256, 0, c->s>>>(c->n, c->prm.eps, dr, 0, dz, 0,
                                                                  nullptr);
        HIPCHK(hipGetLastError());
        return;
    }
    const size_t shm = sizeof(double) * (size_t)c->mmax;
    KTimer kt(c, 4);
    // three waves per leaf: with the kernel's registers a CU holds four such
    // workgroups (three of four waves), and more leaves in flight beat more
    // waves per leaf (measured on the 484k-face surface: 0.467 ms against
    // 0.497 ms with four waves, 1.01 ms with eight)
    if (c->lu32.p)
        k_lu_solve<kSolveNW, float><<<c->nleaf, kSolveNW * kWave, shm, c->s>>>(c->n, c->lorder.p,
            c->blk_ibeg.p, c->blk_m.p, c->blk_off.p, c->piv_off.p, c->lu32.p, c->perm.p, dr, dz, c->gate,
            c->blk_sf.p);
    else
        k_lu_solve<kSolveNW, double><<<c->nleaf, kSolveNW * kWave, shm, c->s>>>(c->n, c->lorder.p,
            c->blk_ibeg.p, c->blk_m.p, c->blk_off.p, c->piv_off.p, c->lu.p, c->perm.p, dr, dz, c->gate,
            c->blk_sf.p);
    HIPCHK(hipGetLastError());
}

double dev_dot(tabi_ctx *c, int64_t n, const double *a, const double *b)
{
    k_dot_part<<<kRedBlocks, 256, 0, c->s>>>(n, a, b, c->red.p);
    k_sum_part<<<1, 256, 0, c->s>>>(kRedBlocks, c->red.p, c->red.p + kRedBlocks);
    double h = 0.0;
    HIPCHK(hipMemcpyAsync(&h, c->red.p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return h;
}

// ---- blocks of vectors ------------------------------------------------------
constexpr int64_t kBlockPartBudget = (int64_t)4 << 30;   // bytes of task partials per pass

template <int P, int RC>
void launch_farb(tabi_ctx *c, int R, int64_t wstride)
{
    const dim3 grid(c->n_far_task, (R + RC - 1) / RC);
    if (c->prm.precision == 1)
        k_far32b<P, RC><<<grid, 64, 0, c->s>>>(c->far_tasks.p, c->n_far_task, c->n, R,
            c->far_node.p, c->far_mask.p, c->node2mom.p, c->mom_wb32.p, wstride, c->ngeo.p,
            c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, (float)c->prm.kappa,
            c->part_farb.p);
    else
        k_farb<P, RC><<<grid, 64, 0, c->s>>>(c->far_tasks.p, c->n_far_task, c->n, R,
            c->far_node.p, c->far_mask.p, c->node2mom.p, c->mom_wb.p, wstride, c->ngeo.p,
            c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, c->prm.kappa,
            c->part_farb.p);
}

template <int P>
void block_fields(tabi_ctx *c, int R, const double *X, int64_t xs, const int32_t *d_act)
{
    constexpr int NK = MomShape<P>::NK, NW = MomShape<P>::NW;
    // moments of every column
    const int64_t pstride = (int64_t)c->nchunk * 4 * NK, wstride = (int64_t)c->nmom * 8 * NW;
    if (c->nmom > 0 && c->n_far_task > 0) {
        KTimer kt(c, 0);
        c->mom_partb.alloc((size_t)R * pstride);
        c->mom_wb.alloc((size_t)R * wstride);
        k_mom_rawb<P><<<dim3(c->nchunk, R), 256, 0, c->s>>>(c->n, c->chunk_node.p, c->chunk_beg.p,
            c->chunk_end.p, c->ngeo.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p,
            c->tarea.p, X, xs, d_act, c->mom_partb.p, pstride);
        k_mom_finalb<P><<<dim3(c->nmom, R), 256, 0, c->s>>>(c->mom_chunk_off.p, c->mom_partb.p,
            pstride, c->prm.eps, c->mom_wb.p, wstride);
        if (c->prm.precision == 1) {
            c->mom_wb32.alloc((size_t)R * wstride);
            k_to_float<<<nblk((int64_t)R * wstride, 256), 256, 0, c->s>>>((int64_t)R * wstride,
                c->mom_wb.p, c->mom_wb32.p);
        }
    }
    // near field
    constexpr int RCN = 8;
    const int Rp = (R + RCN - 1) / RCN * RCN;
    if (c->n_near_task) {
        KTimer kt(c, 1);
        const dim3 gw(nblk(c->n, 256), Rp), gn(c->n_near_task, Rp / RCN);
        if (c->prm.precision == 1) {
            c->wbf.alloc((size_t)c->n * 2 * Rp);
            k_src_weightsb<float><<<gw, 256, 0, c->s>>>(c->n, R, Rp, c->tarea.p, X, xs, d_act, c->wbf.p);
            k_near32b<RCN><<<gn, 64, 0, c->s>>>(c->near_tasks.p, c->near_tgt.p, c->n_near_task, c->n, R, Rp,
                c->src32.p, c->wbf.p, c->ngeo.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p,
                c->tnz.p, (float)c->prm.kappa, (float)c->prm.eps, c->part_nearb.p);
        } else {
            c->wbd.alloc((size_t)c->n * 2 * Rp);
            k_src_weightsb<double><<<gw, 256, 0, c->s>>>(c->n, R, Rp, c->tarea.p, X, xs, d_act, c->wbd.p);
            k_nearb<RCN><<<gn, 64, 0, c->s>>>(c->near_tasks.p, c->near_tgt.p, c->n_near_task, c->n, R, Rp,
                c->src8.p, c->wbd.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p,
                c->prm.kappa, c->prm.eps, c->part_nearb.p);
        }
    }
    // far field: columns per workgroup chosen so the expansion and the
    // accumulators stay in registers
    if (c->n_far_task) {
        KTimer kt(c, 2);
        const bool f32 = c->prm.precision == 1;
        if (P <= 3 || (f32 && P <= 4)) launch_farb<P, 8>(c, R, wstride);
        else if (P == 4 || f32) launch_farb<P, 4>(c, R, wstride);
        else launch_farb<P, 2>(c, R, wstride);
    }
}

// Y[act[col]] = A X[act[col]] for the R block columns (vectors at X + k*xs,
// results at Y + k*ys); columns are processed in passes that bound the task
// partial buffers
void matvec_block_dev(tabi_ctx *c, int R, const double *X, int64_t xs, double *Y, int64_t ys,
                      const int32_t *d_act)
{
    const int64_t per_col = ((int64_t)c->n_near_task + c->n_far_task) * kWave * (int64_t)sizeof(double2);
    const int Rpass = (int)std::max<int64_t>(8, std::min<int64_t>(R, kBlockPartBudget / std::max<int64_t>(per_col, 1)));
    const double pre1 = 0.5 * (1.0 + c->prm.eps), pre2 = 0.5 * (1.0 + 1.0 / c->prm.eps);
    for (int c0 = 0; c0 < R; c0 += Rpass) {
        const int Rc = std::min(Rpass, R - c0);
        const int32_t *act = d_act + c0;
        c->part_nearb.alloc((size_t)std::max(c->n_near_task, 1) * Rc * kWave);
        c->part_farb.alloc((size_t)std::max(c->n_far_task, 1) * Rc * kWave);
        switch (c->p) {
        case 1: block_fields<1>(c, Rc, X, xs, act); break;
        case 2: block_fields<2>(c, Rc, X, xs, act); break;
        case 3: block_fields<3>(c, Rc, X, xs, act); break;
        case 4: block_fields<4>(c, Rc, X, xs, act); break;
        case 5: block_fields<5>(c, Rc, X, xs, act); break;
        case 6: block_fields<6>(c, Rc, X, xs, act); break;
        default: throw Error("unsupported Taylor order");
        }
        KTimer kt(c, 3);
        k_reduceb<<<dim3(nblk(c->n, 256), Rc), 256, 0, c->s>>>(c->n, Rc, near_slots(c),
            c->far_task_off.p, c->part_nearb.p, c->part_farb.p, X, xs, Y, ys, act, pre1, pre2);
    }
    HIPCHK(hipGetLastError());
}

template <int RC>
void launch_lu_solveb(tabi_ctx *c, int R, const double *Rv, int64_t rs, double *Z, int64_t zs,
                      const int32_t *d_act, int ldz)
{
    const size_t shm = sizeof(double) * (size_t)RC * ldz;
    const dim3 grid(c->nleaf, (R + RC - 1) / RC);
    if (c->lu32.p)
        k_lu_solveb<RC, float><<<grid, kWave, shm, c->s>>>(c->n, R, c->blk_ibeg.p, c->blk_m.p,
            c->blk_off.p, c->piv_off.p, c->lu32.p, c->perm.p, Rv, rs, Z, zs, d_act, ldz);
    else
        k_lu_solveb<RC, double><<<grid, kWave, shm, c->s>>>(c->n, R, c->blk_ibeg.p, c->blk_m.p,
            c->blk_off.p, c->piv_off.p, c->lu.p, c->perm.p, Rv, rs, Z, zs, d_act, ldz);
}

void msolve_block_dev(tabi_ctx *c, int R, const double *Rv, int64_t rs, double *Z, int64_t zs,
                      const int32_t *d_act)
{
    KTimer kt(c, 4);
    if (c->prm.flags & 4) {
        k_diag_prec<<<dim3(nblk(2 * (int64_t)c->n, 256), (unsigned)R), 256, 0, c->s>>>(c->n,
            c->prm.eps, Rv, rs, Z, zs, d_act);
        HIPCHK(hipGetLastError());
        return;
    }
    const int ldz = c->mmax;
    const size_t lim = 64 * 1024;
    if (4 * sizeof(double) * (size_t)ldz <= lim) launch_lu_solveb<4>(c, R, Rv, rs, Z, zs, d_act, ldz);
    else if (2 * sizeof(double) * (size_t)ldz <= lim) launch_lu_solveb<2>(c, R, Rv, rs, Z, zs, d_act, ldz);
    else launch_lu_solveb<1>(c, R, Rv, rs, Z, zs, d_act, ldz);
    HIPCHK(hipGetLastError());
}

// Pinned host staging of one host thread for the set-up's bulk transfers, one
// buffer per direction, grown on demand and kept; callers finish every copy
// that uses a buffer (stream synchronisation) before it is asked for again.
struct PinnedBuf {
    char *p = nullptr;
    size_t cap = 0;
    char *get(size_t bytes)
    {
        if (bytes > cap) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&p), bytes, hipHostMallocDefault));
            cap = bytes;
        }
        return p;
    }
};
PinnedBuf &pinned_down() { thread_local PinnedBuf b; return b; }
PinnedBuf &pinned_up() { thread_local PinnedBuf b; return b; }

// Bump allocation in a pinned buffer, 256-byte aligned.
struct PinnedArena {
    char *base;
    size_t used = 0;
    template <class T>
    T *take(size_t count)
    {
        T *r = reinterpret_cast<T *>(base + used);
        used += (count * sizeof(T) + 255) / 256 * 256;
        return r;
    }
};
inline size_t arena_bytes(size_t bytes) { return (bytes + 255) / 256 * 256; }

void build_lists(tabi_ctx *c)
{

    const int G = c->ngroups;
    const HostTree &T = c->tree;
    const int nn = T.nnode();
    c->cnt_far.alloc(G);
    c->cnt_near.alloc(G);
    c->need_mom.alloc(nn);
    c->err_flag.alloc(1);
    HIPCHK(hipMemsetAsync(c->need_mom.p, 0, sizeof(int32_t) * nn, c->s));
    HIPCHK(hipMemsetAsync(c->err_flag.p, 0, sizeof(int32_t), c->s));
    const unsigned nb = nblk(G, 4);
    k_lists<<<nb, 256, 0, c->s>>>(0, c->n, G, c->tx.p, c->ty.p, c->tz.p, c->ngeo.p, c->nnum.p,
        c->nchild.p, c->ncoff.p, c->nkids.p, c->prm.theta, c->cnt_far.p, c->cnt_near.p,
        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->err_flag.p, c->groot.p,
        c->gend.p);
    HIPCHK(hipGetLastError());
    std::vector<int32_t> cf(G), cn(G);
    int32_t ef = 0;
    HIPCHK(hipMemcpyAsync(cf.data(), c->cnt_far.p, sizeof(int32_t) * G, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(cn.data(), c->cnt_near.p, sizeof(int32_t) * G, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(&ef, c->err_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (ef) throw Error("tree traversal stack overflow");
    c->h_far_off.assign(G + 1, 0);
    c->h_near_off.assign(G + 1, 0);
    for (int g = 0; g < G; ++g) {
        c->h_far_off[g + 1] = c->h_far_off[g] + cf[g];
        c->h_near_off[g + 1] = c->h_near_off[g] + cn[g];
    }
    const int64_t nf = c->h_far_off[G], nnr = c->h_near_off[G];
    c->n_far_items = nf;
    c->n_near_items = nnr;
    c->far_off.upload(c->h_far_off.data(), G + 1, c->s);
    c->near_off.upload(c->h_near_off.data(), G + 1, c->s);
    c->far_node.alloc(std::max<int64_t>(nf, 1));
    c->far_mask.alloc(std::max<int64_t>(nf, 1));
    c->near_leaf.alloc(std::max<int64_t>(nnr, 1));
    c->near_mask.alloc(std::max<int64_t>(nnr, 1));
    k_lists<<<nb, 256, 0, c->s>>>(1, c->n, G, c->tx.p, c->ty.p, c->tz.p, c->ngeo.p, c->nnum.p,
        c->nchild.p, c->ncoff.p, c->nkids.p, c->prm.theta, nullptr, nullptr, c->far_off.p,
        c->near_off.p, c->far_node.p, c->far_mask.p, c->near_leaf.p, c->near_mask.p,
        c->need_mom.p, c->err_flag.p, c->groot.p, c->gend.p);
    HIPCHK(hipGetLastError());

    // host: tasks, moment chunks, counters (the far masks stay on the device:
    // only their bit count is needed here)
    DBuf<unsigned long long> dpf;
    dpf.alloc(1);
    HIPCHK(hipMemsetAsync(dpf.p, 0, sizeof(unsigned long long), c->s));
    if (nf) k_popc_sum<<<std::min<unsigned>(nblk(nf, 256), 1024), 256, 0, c->s>>>(nf, c->far_mask.p, dpf.p);
    HIPCHK(hipGetLastError());
    PinnedArena down{pinned_down().get(arena_bytes(sizeof(int32_t) * nnr) + arena_bytes(sizeof(u64) * nnr) +
                                       arena_bytes(sizeof(int32_t) * nn) + 256)};
    int32_t *near_leaf = down.take<int32_t>(nnr);
    u64 *near_mask = down.take<u64>(nnr);
    int32_t *need = down.take<int32_t>(nn);
    unsigned long long *pf = down.take<unsigned long long>(1);
    HIPCHK(hipMemcpyAsync(near_leaf, c->near_leaf.p, sizeof(int32_t) * nnr, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(near_mask, c->near_mask.p, sizeof(u64) * nnr, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(need, c->need_mom.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(pf, dpf.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->p_far = (int64_t)*pf;
    // near field, compacted per leaf (the layout NearSlots describes): the
    // targets that list a leaf are cut into waves of 64, so no lane idles on a
    // target that does not need the leaf
    std::vector<int32_t> lcnt(nn, 0), ipos(std::max<int64_t>(nnr, 1));
    c->p_near = 0;
    for (int64_t k = 0; k < nnr; ++k) {
        const int nd = near_leaf[k];
        const int pc = __builtin_popcountll(near_mask[k]);
        ipos[k] = lcnt[nd];
        lcnt[nd] += pc;
        c->p_near += (int64_t)pc * T.numpar[nd];
    }
    c->p_near -= c->nreal();  // self pairs are excluded

    // task granularity: enough waves to fill the device several times over
    // work per task: the near source runs and far item runs are cut so that
    // each launch has about `target` tasks (flags bits 8-11 override the
    // default of 2^16 with 2^(10+t), t = 1..15).  Packed surfaces are cut
    // each by its own totals, as it would be alone, so every target sums the
    // same task partials in the same order as in its own context.
    const int tt = (c->prm.flags >> 8) & 15;
    const int64_t target = tt ? (int64_t)1 << (10 + tt) : 65536;
    const int nsf = c->nsurf();
    std::vector<int32_t> node0 = c->sf_node0, group0 = c->sf_group0;
    if (c->sf_n.empty()) { node0 = {0, nn}; group0 = {0, G}; }
    std::vector<int32_t> toff(nn), nsplit(nn);
    std::vector<NearTask> nt;
    std::vector<FarTask> ft;
    std::vector<int32_t> fto(G + 1, 0);
    for (int sf = 0; sf < nsf; ++sf) {
        int64_t near_src_total = 0;   // source sweeps of the compacted waves
        for (int nd = node0[sf]; nd < node0[sf + 1]; ++nd)
            if (lcnt[nd]) near_src_total += (int64_t)((lcnt[nd] + kWave - 1) / kWave) * T.numpar[nd];
        const int64_t nfs = c->h_far_off[group0[sf + 1]] - c->h_far_off[group0[sf]];
        int S = 64;
        while (S < 2048 && near_src_total / S > target) S *= 2;
        int F = 1;
        while (F < 64 && nfs / F > target) F *= 2;
        // tasks per leaf: waves of 64 targets x source runs of S
        for (int nd = node0[sf]; nd < node0[sf + 1]; ++nd) {
            if ((int64_t)nt.size() * kWave > INT32_MAX) throw Error("too many near-field tasks");
            toff[nd] = (int32_t)nt.size();
            nsplit[nd] = (T.numpar[nd] + S - 1) / S;
            const int b = T.ibeg[nd], e = T.iend[nd] + 1;
            for (int c0 = 0; c0 < lcnt[nd]; c0 += kWave)
                for (int s0 = b; s0 < e; s0 += S) nt.push_back({s0, std::min(e, s0 + S), nd, sf});
        }
        for (int g = group0[sf]; g < group0[sf + 1]; ++g) {
            for (int32_t k = c->h_far_off[g]; k < c->h_far_off[g + 1]; k += F)
                ft.push_back({g, k, std::min(c->h_far_off[g + 1], k + F), sf});
            fto[g + 1] = (int32_t)ft.size();
        }
    }
    const int64_t ntask = (int64_t)nt.size();
    if (ntask * kWave > INT32_MAX) throw Error("too many near-field tasks");
    c->n_near_task = (int32_t)ntask;
    c->n_far_task = (int32_t)ft.size();

    // moment nodes and their particle chunks
    std::vector<int32_t> n2m(nn, -1), cnode, cbeg, cend, mco(1, 0), msf;
    int32_t nm = 0;
    c->p_mom = 0;
    // surface of each node and group (all 0 for a context of one surface)
    std::vector<int32_t> ndsf(nn, 0), gsf(G, 0);
    for (int sf = 0; sf < nsf; ++sf) {
        for (int nd = node0[sf]; nd < node0[sf + 1]; ++nd) ndsf[nd] = sf;
        for (int g = group0[sf]; g < group0[sf + 1]; ++g) gsf[g] = sf;
    }
    for (int i = 0; i < nn; ++i) {
        if (!need[i]) continue;
        n2m[i] = nm++;
        msf.push_back(ndsf[i]);
        c->p_mom += T.numpar[i];
        for (int b = T.ibeg[i]; b <= T.iend[i]; b += kMomChunk) {
            cnode.push_back(i);
            cbeg.push_back(b);
            cend.push_back(std::min(T.iend[i] + 1, b + kMomChunk));
        }
        mco.push_back((int32_t)cnode.size());
    }
    c->nmom = nm;
    c->nchunk = (int32_t)cnode.size();
    std::vector<int32_t> csf(cnode.size());
    for (size_t k = 0; k < cnode.size(); ++k) csf[k] = ndsf[cnode[k]];

    // every table goes up through one pinned arena, asynchronously
    const bool geom32 = c->prm.precision == 1;
    const size_t up_bytes = arena_bytes(sizeof(int32_t) * ipos.size()) + 2 * arena_bytes(sizeof(int32_t) * nn) +
                            arena_bytes(sizeof(NearTask) * nt.size()) + arena_bytes(sizeof(FarTask) * ft.size()) +
                            arena_bytes(sizeof(int32_t) * fto.size()) + arena_bytes(sizeof(int32_t) * nn) +
                            3 * arena_bytes(sizeof(int32_t) * cnode.size()) + arena_bytes(sizeof(int32_t) * mco.size()) +
                            arena_bytes(sizeof(int32_t) * T.leaves.size()) + arena_bytes(sizeof(int32_t) * G) +
                            arena_bytes(sizeof(int32_t) * csf.size()) + arena_bytes(sizeof(int32_t) * msf.size()) +
                            256;
    PinnedArena up{pinned_up().get(up_bytes)};
    auto put = [&](auto &d, const auto &v) {
        using E = typename std::decay_t<decltype(v)>::value_type;
        E *stage = up.take<E>(v.size());
        std::copy(v.begin(), v.end(), stage);
        d.alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty())
            HIPCHK(hipMemcpyAsync(d.p, stage, sizeof(E) * v.size(), hipMemcpyHostToDevice, c->s));
    };
    c->near_tasks.alloc(std::max<int64_t>(ntask, 1));
    c->far_tasks.alloc(std::max<size_t>(ft.size(), 1));
    put(c->near_ipos, ipos);
    put(c->node_toff, toff);
    put(c->node_nsplit, nsplit);
    if (ntask) {
        put(c->near_tasks, nt);
        c->near_tgt.alloc(ntask * kWave);
        k_fill_i32<<<nblk(ntask * kWave, 256), 256, 0, c->s>>>(ntask * kWave, -1, c->near_tgt.p);
        k_near_tgt<<<G, kWave, 0, c->s>>>(near_slots(c), c->near_tgt.p);
        HIPCHK(hipGetLastError());
    }
    if (!ft.empty()) put(c->far_tasks, ft);
    put(c->far_task_off, fto);
    c->part_near.alloc((size_t)std::max(c->n_near_task, 1) * kWave);
    c->part_far.alloc((size_t)std::max(c->n_far_task, 1) * kWave);
    put(c->node2mom, n2m);
    put(c->group_sf, gsf);
    put(c->chunk_sf, csf);
    put(c->mom_sf, msf);
    if (nm) {
        put(c->chunk_node, cnode);
        put(c->chunk_beg, cbeg);
        put(c->chunk_end, cend);
        put(c->mom_chunk_off, mco);
        c->mom_part.alloc((size_t)c->nchunk * 4 * c->nk);
        c->mom_w.alloc((size_t)nm * 8 * c->nw);
        if (geom32) c->mom_w32.alloc((size_t)nm * 8 * c->nw);
    }
    DBuf<int32_t> dleaves, dfl;
    if (geom32) {
        // leaf-local FP32 source records: each particle's leaf node, filled on the GPU
        put(dleaves, T.leaves);
        dfl.alloc(c->n);
        if (!c->sf_n.empty())   // padding faces of packed surfaces: any valid node
            k_fill_i32<<<nblk(c->n, 256), 256, 0, c->s>>>(c->n, 0, dfl.p);
        k_leaf_fill<<<(unsigned)T.leaves.size(), 256, 0, c->s>>>(dleaves.p, c->nibeg.p, c->nnum.p, dfl.p);
        c->src32.alloc((size_t)c->n * 8);
        k_src32_geom<<<nblk(c->n, 256), 256, 0, c->s>>>(c->n, dfl.p, c->ngeo.p, c->tx.p, c->ty.p,
            c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, c->src32.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->s));
}

// Workgroups are dealt to the 8 XCDs round-robin in dispatch order.  A tile
// list is laid out so that XCD x receives the contiguous run
// [x*per, (x+1)*per) of the logical order (per = ceil(count/8)): tiles of
// one leaf block, which share their L and U panels, then meet in one L2.
// Padding entries carry block -1 and exit at once.  Lists of at most 8 tiles
// keep the logical order.
constexpr int kXcd = 8;

// The tile lists of the batched LU, one entry per workgroup of a launch.
enum LuTileKind : int {
    kTilesSwap = 0,    // {block, column tile}: in-strip solve right of the panel
    kTilesPanel = 1,   // {block, tile row, tile col, strip end}: update after a panel
    kTilesStrip = 2,   // {block, column tile}: solve of the columns beyond a strip
    kTilesBelow = 3,   // {block, tile row, tile col, m}: update below a strip
    kTileKinds = 4
};

// Tiles of one kind that block m contributes at panel step k0 (the logical
// order within a block is row-major over its tile grid).
__host__ __device__ inline void lu_tile_grid(int kind, int k0, int mq, int &ny, int &nx)
{
    const int K0 = k0 - k0 % kLuW, send = mq < K0 + kLuW ? mq : K0 + kLuW;
    const bool strip_end = k0 + kLuNB == K0 + kLuW;
    ny = 1; nx = 0;
    if (kind == kTilesSwap) {
        // the column tiles holding strip columns right of the panel
        if (send > k0 + kLuNB) nx = (send - 1) / 64 - (k0 + kLuNB) / 64 + 1;
    } else if (kind == kTilesPanel) {
        const int rem = mq - k0 - kLuNB, cols = send - k0 - kLuNB;
        if (rem > 0 && cols > 0) { ny = (rem + 63) / 64; nx = (cols + 63) / 64; }
    } else if (strip_end && mq > K0 + kLuW) {
        nx = (mq - K0 - kLuW + 63) / 64;
        ny = kind == kTilesBelow ? nx : 1;
    }
}

// Expands one kind's per-step tile counts into its XCD-grouped list.  Step s
// (k0 = s * kLuNB) owns entries [off[s], off[s+1]) and the exclusive prefix
// pre[poff[s] + i] of the tile counts of the active blocks lorder[0..na[s]);
// an entry's logical index q is mapped back through the XCD layout and its
// block found by bisection.  One thread per entry.
__global__ void k_lu_tiles(int kind, int nsteps, const int64_t *__restrict__ off,
    const int32_t *__restrict__ cnt, const int32_t *__restrict__ poff,
    const int32_t *__restrict__ na, const int32_t *__restrict__ pre,
    const int32_t *__restrict__ lorder, const int32_t *__restrict__ blk_m, int32_t *__restrict__ out)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= off[nsteps]) return;
    int s = 0;
    while (off[s + 1] <= b) ++s;
    const int64_t r = b - off[s];
    const int c = cnt[s];
    int64_t q = r;
    if (c > kXcd) {
        const int64_t per = (c + kXcd - 1) / kXcd;
        q = (r % kXcd) * per + r / kXcd;
    }
    int4 v = make_int4(-1, 0, 0, 0);
    if (q < c) {
        const int32_t *pr = pre + poff[s];
        int lo = 0, hi = na[s];                // last i with pr[i] <= q
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (pr[mid] <= q) lo = mid; else hi = mid;
        }
        const int l = lorder[lo], idx = (int)(q - pr[lo]), mq = blk_m[l];
        const int k0 = s * kLuNB;
        int ny, nx;
        lu_tile_grid(kind, k0, mq, ny, nx);
        const int K0 = k0 - k0 % kLuW;
        if (kind == kTilesPanel) v = make_int4(l, idx / nx, idx % nx, min(mq, K0 + kLuW));
        else if (kind == kTilesBelow) v = make_int4(l, idx / nx, idx % nx, mq);
        else if (kind == kTilesSwap) v = make_int4(l, (k0 + kLuNB) / 64 + idx, 0, 0);
        else v = make_int4(l, idx, 0, 0);
    }
    if (kind == kTilesPanel || kind == kTilesBelow) {
        reinterpret_cast<int4 *>(out)[b] = v;
    } else {
        reinterpret_cast<int2 *>(out)[b] = make_int2(v.x, v.y);
    }
}

// Side streams (default priority) for launches that may run beside the
// library's stream, one set per device, created on first use and kept.  Each
// is created with a mask of every CU, which gives it a hardware queue of its
// own: plain streams share the runtime's few queues, and two streams on one
// queue run one after the other.  (Measured on the 484k-face surface: the
// panel classes' side streams had landed on one queue together, and the
// right-hand side and the learned guess on the library's stream's queue
// behind the factorisation; with their own queues the set-up of a solve went
// 19.7 -> 18.2 ms.)
hipStream_t side_stream(int k)
{
    static std::mutex mu;
    static std::map<int, std::array<hipStream_t, 3>> pool;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto &a = pool[dev];
    if (!a[k]) {
        int ncu = 0;
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
        for (int q = 0; q < ncu; ++q) mask[(size_t)q / 32] |= 1u << (q % 32);
        HIPCHK(hipExtStreamCreateWithCUMask(&a[k], (uint32_t)mask.size(), mask.data()));
    }
    return a[k];
}

template <class T>
void xcd_group(std::vector<T> &v, size_t beg, T pad)
{
    const size_t cnt = v.size() - beg;
    if (cnt <= (size_t)kXcd) return;
    const size_t per = (cnt + kXcd - 1) / kXcd;
    std::vector<T> out(per * kXcd, pad);
    for (size_t q = 0; q < cnt; ++q) out[(q % per) * kXcd + q / per] = v[beg + q];
    v.resize(beg);
    v.insert(v.end(), out.begin(), out.end());
}

// Leaf-block tables of the preconditioner (block sizes and offsets, the face
// -> block map), uploaded on the context's stream; the learned-guess features
// read them too.
void precond_tables(tabi_ctx *c)
{
    const HostTree &T = c->tree;
    const int nl = (int)T.leaves.size();
    c->nleaf = nl;
    std::vector<int32_t> bib(nl), bm(nl), poff(nl), leaf_of(c->n, -1);
    c->h_blk_off.assign(nl, 0);
    int64_t off = 0;
    int32_t pof = 0;
    c->mmax = 0;
    for (int l = 0; l < nl; ++l) {
        const int nd = T.leaves[l];
        bib[l] = T.ibeg[nd];
        bm[l] = 2 * T.numpar[nd];
        c->h_blk_off[l] = off;
        off += (int64_t)bm[l] * bm[l];
        poff[l] = pof;
        pof += bm[l];
        c->mmax = std::max(c->mmax, bm[l]);
        for (int i = T.ibeg[nd]; i <= T.iend[nd]; ++i) leaf_of[i] = l;
    }
    c->h_blk_m = bm;
    {
        std::vector<int32_t> bsf(nl, 0);
        if (!c->sf_n.empty())
            for (int l = 0; l < nl; ++l)
                bsf[l] = (int32_t)(std::upper_bound(c->sf_node0.begin(), c->sf_node0.end(), T.leaves[l]) -
                                   c->sf_node0.begin()) - 1;
        c->blk_sf.upload(bsf.data(), nl, c->s);
    }
    c->blk_ibeg.upload(bib.data(), nl, c->s);
    c->blk_m.upload(bm.data(), nl, c->s);
    c->blk_off.upload(c->h_blk_off.data(), nl, c->s);
    c->piv_off.upload(poff.data(), nl, c->s);
    c->leaf_of.upload(leaf_of.data(), c->n, c->s);
    c->piv.alloc(pof);
    c->lu.alloc(off);
    c->lu_info.alloc(nl);
    c->lu_swrec.alloc((size_t)nl * kSwRec);
    c->lu_linv.alloc((size_t)nl * (kLuW / kLuNB) * kLuNB * kLuNB);
    HIPCHK(hipMemsetAsync(c->lu_info.p, 0, sizeof(int32_t) * nl, c->s));
}

// Queues the assembly and factorisation of every leaf block (and, for mixed
// precision, the rounding of the factors to FP32) on the context's stream and
// its side streams, without waiting: work on other streams can run beside it
// until precond_finish.
void precond_factor(tabi_ctx *c)
{
    const int nl = c->nleaf;
    const std::vector<int32_t> &bm = c->h_blk_m;
    const int64_t off = (int64_t)c->lu.n;
    // mixed precision: the solves read FP32 factors.  The kernels that finish
    // an entry (the panels, the in-strip and strip-end solves, the
    // interchanges) also write its FP32 rounding; flags bit 5 instead rounds
    // the finished FP64 factors in one pass (the same values)
    const bool mixed = c->prm.precision == 1, one_pass = (c->prm.flags & 32) != 0;
    float *l32 = nullptr;
    if (mixed) {
        c->lu32.alloc(off);
        if (!one_pass) l32 = c->lu32.p;
    }
    // the tile lists below are uploaded on a side stream that waits for the
    // work queued so far but not for the assembly, so the copies overlap it
    c->pre_ev.make(5);
    hipEvent_t ev_fork = c->pre_ev.e[0], ev_up = c->pre_ev.e[1], *ev_join = &c->pre_ev.e[2];
    HIPCHK(hipEventRecord(ev_up, c->s));
    k_lu_assemble<<<nblk(c->n, 4 * kAsmPerWave), 256, 0, c->s>>>(c->n, c->leaf_of.p, c->blk_ibeg.p, c->blk_m.p,
        c->blk_off.p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, c->tarea.p,
        c->prm.kappa, c->prm.eps, c->prm.flags, c->lu.p);
    HIPCHK(hipGetLastError());
    // largest blocks first: the active blocks of each panel step are a prefix
    std::vector<int32_t> lorder(nl);
    for (int l = 0; l < nl; ++l) lorder[l] = l;
    std::stable_sort(lorder.begin(), lorder.end(), [&](int a, int b) { return bm[a] > bm[b]; });
    // per-step tile counts of each list kind; the lists themselves are
    // expanded on the GPU (k_lu_tiles) beside the assembly
    const int nst = (c->mmax + kLuNB - 1) / kLuNB;
    std::vector<int> nact(nst);
    for (int st = 0; st < nst; ++st) {
        int na = st ? nact[st - 1] : nl;
        while (na > 0 && bm[lorder[na - 1]] <= st * kLuNB) --na;
        nact[st] = na;
    }
    // tab: per kind cnt[nst] and poff[nst], then na[nst], then the prefixes
    std::vector<int32_t> tab((size_t)(2 * kTileKinds + 1) * nst);
    std::vector<int64_t> toff((size_t)kTileKinds * (nst + 1), 0);
    for (int st = 0; st < nst; ++st) tab[(size_t)2 * kTileKinds * nst + st] = nact[st];
    for (int kind = 0; kind < kTileKinds; ++kind) {
        int64_t *o = toff.data() + (size_t)kind * (nst + 1);
        for (int st = 0; st < nst; ++st) {
            tab[(size_t)(2 * kind + 1) * nst + st] = (int32_t)tab.size();
            int64_t run = 0;
            for (int q = 0; q < nact[st]; ++q) {
                tab.push_back((int32_t)run);
                int ny, nx;
                lu_tile_grid(kind, st * kLuNB, bm[lorder[q]], ny, nx);
                run += (int64_t)ny * nx;
            }
            if (run > INT32_MAX) throw Error("too many LU tiles in one step");
            tab[(size_t)2 * kind * nst + st] = (int32_t)run;
            o[st + 1] = o[st] + (run > kXcd ? (run + kXcd - 1) / kXcd * kXcd : run);
        }
    }
    const int64_t *sw_off = toff.data(), *tl_off = sw_off + (nst + 1);
    const int64_t *ts_off = tl_off + (nst + 1), *tlo_off = ts_off + (nst + 1);
    // the tile lists outlive this call (the queued kernels read them)
    DBuf<int4> &dtl = c->lu_tl, &dtlo = c->lu_tlo;
    DBuf<int2> &dsw = c->lu_sw, &dts = c->lu_ts;
    DBuf<int32_t> &dtab = c->lu_tab;
    DBuf<int64_t> &dtoff = c->lu_toff;
    {
        const hipStream_t up = side_stream(0);
        HIPCHK(hipStreamWaitEvent(up, ev_up, 0));
        c->lorder.upload(lorder.data(), nl, up);
        dtab.upload(tab.data(), tab.size(), up);
        dtoff.upload(toff.data(), toff.size(), up);
        dsw.alloc(std::max<int64_t>(sw_off[nst], 1));
        dtl.alloc(std::max<int64_t>(tl_off[nst], 1));
        dts.alloc(std::max<int64_t>(ts_off[nst], 1));
        dtlo.alloc(std::max<int64_t>(tlo_off[nst], 1));
        int32_t *outs[kTileKinds] = {reinterpret_cast<int32_t *>(dsw.p), reinterpret_cast<int32_t *>(dtl.p),
                                     reinterpret_cast<int32_t *>(dts.p), reinterpret_cast<int32_t *>(dtlo.p)};
        for (int kind = 0; kind < kTileKinds; ++kind) {
            const int64_t tot = toff[(size_t)kind * (nst + 1) + nst];
            if (tot == 0) continue;
            k_lu_tiles<<<(unsigned)((tot + 255) / 256), 256, 0, up>>>(kind, nst,
                dtoff.p + (size_t)kind * (nst + 1), dtab.p + (size_t)2 * kind * nst,
                dtab.p + (size_t)(2 * kind + 1) * nst, dtab.p + (size_t)2 * kTileKinds * nst, dtab.p,
                c->lorder.p, c->blk_m.p, outs[kind]);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev_up, up));
        HIPCHK(hipStreamWaitEvent(c->s, ev_up, 0));
    }
    const int32_t *dlorder_p = c->lorder.p;
    // the panel launches of one step (one per size class) run side by side:
    // the first on the library's stream, the others on side streams that
    // fork from and join back into it
    int step = 0;
    for (int k0 = 0; k0 < c->mmax; k0 += kLuNB, ++step) {
        const int na = nact[step];
        // one launch per size class of the remaining rows (lorder is sorted, so
        // each class is a contiguous range of active blocks)
        int q = 0, ncls = 0;
        HIPCHK(hipEventRecord(ev_fork, c->s));
        while (q < na) {
            const int rem = bm[lorder[q]] - k0;
            int cls = rem > 1024 ? 2048 : rem > 512 ? 1024 : rem > 256 ? 512 : 256;
            int e = q;
            const int lo = cls == 256 ? 0 : cls / 2;
            while (e < na && bm[lorder[e]] - k0 > lo) ++e;
            const int cnt = e - q;
            hipStream_t ps = c->s;
            if (ncls > 0) {
                ps = side_stream(ncls - 1);
                HIPCHK(hipStreamWaitEvent(ps, ev_fork, 0));
            }
            if (cls == 2048)
                k_lu_panel<2, 1024><<<cnt, 1024, 0, ps>>>(k0, q, dlorder_p, c->blk_m.p,
                    c->blk_off.p, c->piv_off.p, c->lu.p, c->piv.p, c->lu_info.p, c->lu_swrec.p, l32);
            else if (cls == 1024)
                k_lu_panel<1, 1024><<<cnt, 1024, 0, ps>>>(k0, q, dlorder_p, c->blk_m.p,
                    c->blk_off.p, c->piv_off.p, c->lu.p, c->piv.p, c->lu_info.p, c->lu_swrec.p, l32);
            else if (cls == 512)
                k_lu_panel<1, 512><<<cnt, 512, 0, ps>>>(k0, q, dlorder_p, c->blk_m.p,
                    c->blk_off.p, c->piv_off.p, c->lu.p, c->piv.p, c->lu_info.p, c->lu_swrec.p, l32);
            else
                k_lu_panel<1, 256><<<cnt, 256, 0, ps>>>(k0, q, dlorder_p, c->blk_m.p,
                    c->blk_off.p, c->piv_off.p, c->lu.p, c->piv.p, c->lu_info.p, c->lu_swrec.p, l32);
            // the class's panel inverses follow its panels on the same stream,
            // beside the other classes' panels
            k_lu_linv<<<cnt, 64, 0, ps>>>(k0, q, dlorder_p, c->blk_m.p, c->blk_off.p, c->lu.p,
                                          c->lu_linv.p);
            if (ncls > 0) HIPCHK(hipEventRecord(ev_join[ncls - 1], ps));
            ++ncls;
            q = e;
        }
        for (int k = 1; k < ncls; ++k) HIPCHK(hipStreamWaitEvent(c->s, ev_join[k - 1], 0));
        const int64_t nsw = sw_off[step + 1] - sw_off[step];
        if (nsw > 0)
            k_lu_swap_trsm<<<(unsigned)nsw, 256, 0, c->s>>>(k0, dsw.p + sw_off[step], c->blk_m.p,
                c->blk_off.p, c->lu.p, c->lu_swrec.p, c->lu_linv.p, l32);
        const int64_t nt = tl_off[step + 1] - tl_off[step];
        if (nt > 0)
            k_lu_gemm<kLuNB><<<(unsigned)nt, 256, 0, c->s>>>(k0, dtl.p + tl_off[step], c->blk_m.p,
                                                             c->blk_off.p, c->lu.p);
        const int64_t nts = ts_off[step + 1] - ts_off[step];
        if (nts > 0)
            k_lu_strip_trsm<<<(unsigned)nts, 256, 0, c->s>>>(k0 + kLuNB - kLuW, dts.p + ts_off[step],
                c->blk_m.p, c->blk_off.p, c->lu.p, c->lu_linv.p, l32);
        const int64_t nto = tlo_off[step + 1] - tlo_off[step];
        if (nto > 0)
            k_lu_gemm<kLuW><<<(unsigned)nto, 256, 0, c->s>>>(k0 + kLuNB - kLuW, dtlo.p + tlo_off[step],
                                                             c->blk_m.p, c->blk_off.p, c->lu.p);
    }
    HIPCHK(hipGetLastError());
    c->perm.alloc(c->piv.n);
    k_lu_perm<<<nl, kWave, 0, c->s>>>(c->blk_m.p, c->piv_off.p, c->piv.p, c->perm.p);
    HIPCHK(hipGetLastError());
    if (mixed && one_pass) {
        k_to_float<<<to_float_grid(off), 256, 0, c->s>>>(off, c->lu.p, c->lu32.p);
        HIPCHK(hipGetLastError());
    }
}

// Waits for the factorisation, reports a singular leaf block (its first zero
// pivot) and frees what only the factorisation used.
void precond_finish(tabi_ctx *c)
{
    const int nl = c->nleaf;
    std::vector<int32_t> info(nl);
    HIPCHK(hipMemcpyAsync(info.data(), c->lu_info.p, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->pre_ev.clear();
    c->lu_tl.release(); c->lu_tlo.release(); c->lu_sw.release(); c->lu_ts.release();
    c->lu_tab.release(); c->lu_toff.release();
    if (c->prm.precision == 1) c->lu.release();
    for (int l = 0; l < nl; ++l)
        if (info[l]) throw Error("singular leaf block " + std::to_string(l + 1) + " (column " +
                                 std::to_string(info[l]) + ")");
}

void build_precond(tabi_ctx *c)
{
    precond_tables(c);
    precond_factor(c);
    precond_finish(c);
}

// Centroids from interleaved {x,y,z} to three arrays.
__global__ void k_deinterleave(int n, const double *__restrict__ xyz, double *__restrict__ x,
                               double *__restrict__ y, double *__restrict__ z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = xyz[3 * i]; y[i] = xyz[3 * i + 1]; z[i] = xyz[3 * i + 2];
}

// Normals and areas into tree order, and the source records {x,y,z,nx,ny,nz,
// 0,0} (the weights are written per matvec).
__global__ void k_gather_geom(int n, const int32_t *__restrict__ order,
                              const double *__restrict__ nrm, const double *__restrict__ area,
                              const double *__restrict__ tx, const double *__restrict__ ty,
                              const double *__restrict__ tz, double *__restrict__ tnx,
                              double *__restrict__ tny, double *__restrict__ tnz,
                              double *__restrict__ tarea, double *__restrict__ src8)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = order[i];
    const double nx = nrm[3 * o], ny = nrm[3 * o + 1], nz = nrm[3 * o + 2];
    tnx[i] = nx; tny[i] = ny; tnz[i] = nz; tarea[i] = area[o];
    double4 *r = reinterpret_cast<double4 *>(src8 + (size_t)8 * i);
    r[0] = make_double4(tx[i], ty[i], tz[i], nx);
    r[1] = make_double4(ny, nz, 0.0, 0.0);
}

void ctx_setup(tabi_ctx *c, const tabi_params *prm, int32_t n, const double *xyz,
               const double *nrm, const double *area, int32_t *orderarr_out, double *t_phase,
               bool dev = false, bool factor = true)
{
    if (!prm) throw Error("null parameters");
    if (n < 1) throw Error("need at least one face");
    if (prm->order < 1 || prm->order > kMaxOrder) throw Error("Taylor order must be 1..6");
    if (prm->maxparnode < 1) throw Error("maxparnode must be >= 1");
    if (prm->precision != 0 && prm->precision != 1) throw Error("precision must be 0 or 1");
    c->prm = *prm;
    c->n = n;
    c->p = prm->order;
    c->nk = (c->p + 1) * (c->p + 2) * (c->p + 3) / 6;
    c->nw = (c->p + 2) * (c->p + 3) * (c->p + 4) / 6;
    c->ngroups = (n + kWave - 1) / kWave;
    if (!c->s) HIPCHK(hipStreamCreate(&c->s));
    PoolScope ps_(c->s);
    auto t0 = std::chrono::steady_clock::now();
    // inputs in original face order, host or device
    DevIn d_xyz(xyz, (size_t)3 * n, dev, c->s), d_nrm(nrm, (size_t)3 * n, dev, c->s),
        d_area(area, n, dev, c->s);
    DBuf<int32_t> d_ord;
    c->tx.alloc(n); c->ty.alloc(n); c->tz.alloc(n);
    d_ord.alloc(n);
    if (prm->flags & 2) {
        std::vector<double> hxyz;
        if (dev) {
            hxyz.resize((size_t)3 * n);
            HIPCHK(hipMemcpyAsync(hxyz.data(), xyz, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->s));
            HIPCHK(hipStreamSynchronize(c->s));
            xyz = hxyz.data();
        }
        std::vector<double> x(n), y(n), z(n);
        for (int i = 0; i < n; ++i) { x[i] = xyz[3 * i]; y[i] = xyz[3 * i + 1]; z[i] = xyz[3 * i + 2]; }
        tabi::build_tree_host(n, x.data(), y.data(), z.data(), prm->maxparnode, c->tree);
        c->tx.upload(x.data(), n, c->s); c->ty.upload(y.data(), n, c->s); c->tz.upload(z.data(), n, c->s);
        d_ord.upload(c->tree.order.data(), n, c->s);
    } else {
        DBuf<double> dx, dy, dz;
        dx.alloc(n); dy.alloc(n); dz.alloc(n);
        k_deinterleave<<<nblk(n, 256), 256, 0, c->s>>>(n, d_xyz.p, dx.p, dy.p, dz.p);
        HIPCHK(hipGetLastError());
        tabi::build_tree_gpu(n, dx.p, dy.p, dz.p, prm->maxparnode, c->s, c->tree, c->tx.p,
                             c->ty.p, c->tz.p, d_ord.p);
    }
    const HostTree &T = c->tree;
    if (orderarr_out)
        for (int i = 0; i < n; ++i) orderarr_out[i] = T.order[i] + 1;
    c->tnx.alloc(n); c->tny.alloc(n); c->tnz.alloc(n); c->tarea.alloc(n);
    c->src8.alloc((size_t)n * 8);
    k_gather_geom<<<nblk(n, 256), 256, 0, c->s>>>(n, d_ord.p, d_nrm.p, d_area.p, c->tx.p, c->ty.p,
                                                  c->tz.p, c->tnx.p, c->tny.p, c->tnz.p,
                                                  c->tarea.p, c->src8.p);
    HIPCHK(hipGetLastError());
    const int nn = T.nnode();
    std::vector<double4> geo(nn);
    for (int k = 0; k < nn; ++k) geo[k] = make_double4(T.xmid[k], T.ymid[k], T.zmid[k], T.radius[k]);
    c->ngeo.upload(geo.data(), nn, c->s);
    c->nnum.upload(T.numpar.data(), nn, c->s);
    c->nchild.upload(T.nchild.data(), nn, c->s);
    c->ncoff.upload(T.child_off.data(), nn, c->s);
    c->nibeg.upload(T.ibeg.data(), nn, c->s);
    c->nkids.upload(T.children.empty() ? T.nchild.data() : T.children.data(),
                    std::max<size_t>(T.children.size(), 1), c->s);
    c->vx.alloc((size_t)2 * n);
    c->vy.alloc((size_t)2 * n);
    c->red.alloc(kRedBlocks + 8);
    HIPCHK(hipStreamSynchronize(c->s));
    auto t1 = std::chrono::steady_clock::now();
    build_lists(c);
    auto t2 = std::chrono::steady_clock::now();
    // without factor only the leaf-block tables: the caller queues the
    // factorisation (precond_factor) and waits for it (precond_finish)
    if (prm->flags & 4)
        c->nleaf = (int32_t)c->tree.leaves.size();   // diagonal scaling: no leaf blocks
    else if (factor)
        build_precond(c);
    else
        precond_tables(c);
    auto t3 = std::chrono::steady_clock::now();
    if (t_phase) {
        t_phase[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t_phase[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
        t_phase[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
    }
}

// One context over several surfaces ("packed"): surface r's faces occupy
// [F_r, F_r + n_r) in a common numbering, F_r a multiple of 64 so a target group
// of 64 faces never spans two surfaces; the faces between surfaces are
// padding (zero geometry and area, in no leaf and no group's targets).  Each
// surface keeps its own tree (built exactly as alone), merged into one node
// table; a group's interaction lists are built from its own surface's root,
// so no interaction crosses surfaces, and the tasks of each surface are cut by
// its own totals.  The field, moment and leaf-block kernels then serve every
// surface in one launch each, and a surface's results are those of its own
// context bit for bit.  Vectors are [first halves of all surfaces (c->n
// entries); second halves], surface r's at F_r and c->n + F_r.
void ctx_setup_packed(tabi_ctx *c, const tabi_params *prm, int nsurf, const int32_t *nface,
                      const double *const *xyz, const double *const *nrm,
                      const double *const *area, double *t_phase = nullptr)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!prm) throw Error("null parameters");
    if (nsurf < 1) throw Error("need at least one surface");
    if (prm->order < 1 || prm->order > kMaxOrder) throw Error("Taylor order must be 1..6");
    if (prm->maxparnode < 1) throw Error("maxparnode must be >= 1");
    if (prm->precision != 0 && prm->precision != 1) throw Error("precision must be 0 or 1");
    c->prm = *prm;
    c->p = prm->order;
    c->nk = (c->p + 1) * (c->p + 2) * (c->p + 3) / 6;
    c->nw = (c->p + 2) * (c->p + 3) * (c->p + 4) / 6;
    c->sf_face0.assign(nsurf + 1, 0);
    c->sf_n.assign(nface, nface + nsurf);
    int64_t F = 0;
    for (int r = 0; r < nsurf; ++r) {
        if (nface[r] < 1) throw Error("surface " + std::to_string(r) + " has no faces");
        c->sf_face0[r] = (int32_t)F;
        F += (nface[r] + kWave - 1) / kWave * kWave;
        if (F > INT32_MAX / 2) throw Error("packed surfaces too large");
    }
    c->sf_face0[nsurf] = (int32_t)F;
    const int32_t n = (int32_t)F;
    c->n = n;
    c->ngroups = n / kWave;
    if (!c->s) HIPCHK(hipStreamCreate(&c->s));
    PoolScope ps_(c->s);
    c->tx.alloc(n); c->ty.alloc(n); c->tz.alloc(n);
    c->tnx.alloc(n); c->tny.alloc(n); c->tnz.alloc(n); c->tarea.alloc(n);
    c->src8.alloc((size_t)n * 8);
    for (DBuf<double> *b : {&c->tx, &c->ty, &c->tz, &c->tnx, &c->tny, &c->tnz, &c->tarea})
        HIPCHK(hipMemsetAsync(b->p, 0, sizeof(double) * n, c->s));
    HIPCHK(hipMemsetAsync(c->src8.p, 0, sizeof(double) * 8 * (size_t)n, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    // trees, several surfaces at a time from host threads (own streams), each
    // writing its own range of the packed arrays
    std::vector<HostTree> trees(nsurf);
    {
        const int nth = std::max(1, std::min(nsurf, 8));
        std::vector<std::thread> th;
        std::vector<std::string> errs(nth);
        for (int t = 0; t < nth; ++t)
            th.emplace_back([&, t] {
                hipStream_t st = nullptr;
                try {
                    HIPCHK(hipStreamCreate(&st));
                    PoolScope pt_(st);
                    for (int r = t; r < nsurf; r += nth) {
                        const int32_t nr = nface[r];
                        const int64_t f0 = c->sf_face0[r];
                        DevIn d_xyz(xyz[r], (size_t)3 * nr, false, st), d_nrm(nrm[r], (size_t)3 * nr, false, st),
                            d_area(area[r], nr, false, st);
                        DBuf<double> dx, dy, dz;
                        DBuf<int32_t> d_ord;
                        dx.alloc(nr); dy.alloc(nr); dz.alloc(nr); d_ord.alloc(nr);
                        k_deinterleave<<<nblk(nr, 256), 256, 0, st>>>(nr, d_xyz.p, dx.p, dy.p, dz.p);
                        HIPCHK(hipGetLastError());
                        tabi::build_tree_gpu(nr, dx.p, dy.p, dz.p, prm->maxparnode, st, trees[r],
                                             c->tx.p + f0, c->ty.p + f0, c->tz.p + f0, d_ord.p);
                        k_gather_geom<<<nblk(nr, 256), 256, 0, st>>>(nr, d_ord.p, d_nrm.p, d_area.p,
                            c->tx.p + f0, c->ty.p + f0, c->tz.p + f0, c->tnx.p + f0, c->tny.p + f0,
                            c->tnz.p + f0, c->tarea.p + f0, c->src8.p + 8 * f0);
                        HIPCHK(hipGetLastError());
                        HIPCHK(hipStreamSynchronize(st));
                    }
                } catch (const std::exception &e) {
                    errs[t] = e.what();
                }
                destroy_stream(st);
            });
        for (auto &x : th) x.join();
        for (auto &e : errs)
            if (!e.empty()) throw Error(e);
    }
    // one node table: surface r's nodes follow surface r-1's
    HostTree &T = c->tree;
    T = HostTree{};
    T.n = n;
    c->sf_node0.assign(nsurf + 1, 0);
    c->sf_group0.assign(nsurf + 1, 0);
    for (int r = 0; r < nsurf; ++r) {
        const HostTree &S = trees[r];
        const int32_t no = T.nnode(), fo = c->sf_face0[r], co = (int32_t)T.children.size();
        c->sf_node0[r] = no;
        c->sf_group0[r] = fo / kWave;
        for (int k = 0; k < S.nnode(); ++k) {
            T.ibeg.push_back(S.ibeg[k] + fo);
            T.iend.push_back(S.iend[k] + fo);
            T.child_off.push_back(S.child_off[k] + co);
        }
        auto app = [](auto &d, const auto &v) { d.insert(d.end(), v.begin(), v.end()); };
        app(T.numpar, S.numpar); app(T.level, S.level); app(T.nchild, S.nchild);
        app(T.xmin, S.xmin); app(T.xmax, S.xmax); app(T.ymin, S.ymin); app(T.ymax, S.ymax);
        app(T.zmin, S.zmin); app(T.zmax, S.zmax);
        app(T.xmid, S.xmid); app(T.ymid, S.ymid); app(T.zmid, S.zmid); app(T.radius, S.radius);
        for (int32_t k : S.children) T.children.push_back(k + no);
        for (int32_t k : S.leaves) T.leaves.push_back(k + no);
    }
    c->sf_node0[nsurf] = T.nnode();
    c->sf_group0[nsurf] = c->ngroups;
    std::vector<int32_t> groot(c->ngroups), gend(c->ngroups);
    for (int r = 0; r < nsurf; ++r)
        for (int g = c->sf_group0[r]; g < c->sf_group0[r + 1]; ++g) {
            groot[g] = c->sf_node0[r];
            gend[g] = c->sf_face0[r] + nface[r];
        }
    c->groot.upload(groot.data(), groot.size(), c->s);
    c->gend.upload(gend.data(), gend.size(), c->s);
    const int nn = T.nnode();
    std::vector<double4> geo(nn);
    for (int k = 0; k < nn; ++k) geo[k] = make_double4(T.xmid[k], T.ymid[k], T.zmid[k], T.radius[k]);
    c->ngeo.upload(geo.data(), nn, c->s);
    c->nnum.upload(T.numpar.data(), nn, c->s);
    c->nchild.upload(T.nchild.data(), nn, c->s);
    c->ncoff.upload(T.child_off.data(), nn, c->s);
    c->nibeg.upload(T.ibeg.data(), nn, c->s);
    c->nkids.upload(T.children.empty() ? T.nchild.data() : T.children.data(),
                    std::max<size_t>(T.children.size(), 1), c->s);
    c->vx.alloc((size_t)2 * n);
    c->vy.alloc((size_t)2 * n);
    c->red.alloc(kRedBlocks + 8);
    HIPCHK(hipStreamSynchronize(c->s));
    auto t1 = std::chrono::steady_clock::now();
    build_lists(c);
    auto t2 = std::chrono::steady_clock::now();
    if (prm->flags & 4)
        c->nleaf = (int32_t)c->tree.leaves.size();
    else
        build_precond(c);
    auto t3 = std::chrono::steady_clock::now();
    if (t_phase) {
        t_phase[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t_phase[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
        t_phase[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
    }
}

void ctx_stats(tabi_ctx *c, int64_t *o)
{
    o[0] = c->n;
    o[1] = c->tree.nnode();
    o[2] = c->nleaf;
    o[3] = c->p_near;
    o[4] = c->p_far;
    o[5] = c->p_mom;
    o[6] = c->n_far_items;
    o[7] = c->n_near_items;
    o[8] = (int64_t)c->n_near_task + c->n_far_task;
    o[9] = c->lu.n * (int64_t)sizeof(double) + c->lu32.n * (int64_t)sizeof(float);
}


// Waits for an event by polling: the wake-up latency of a blocking wait is
// tens of microseconds, several times the work between the GMRES round trips.
void spin_wait(hipEvent_t ev)
{
    for (;;) {
        const hipError_t q = hipEventQuery(ev);
        if (q == hipSuccess) return;
        if (q != hipErrorNotReady) HIPCHK(q);
    }
}

constexpr int kCoefRing = 4;   // coefficient upload slots

// Pinned host memory for the GMRES scalars of one host thread, kept for the
// life of the thread: val and flag are coherent and mapped (written by
// k_sum_slots, read by the host while kernels run), coef stages uploads.
struct PinnedScalars {
    double *val = nullptr, *coef = nullptr, *dval = nullptr;
    int64_t *flag = nullptr, *dflag = nullptr;
    size_t nval = 0, nflag = 0, ncoef = 0;
    int64_t seq = 0;           // last flag value published
};

PinnedScalars &pinned_scalars(size_t nval, size_t nflag, size_t ncoef)
{
    thread_local PinnedScalars ps;
    if (nval > ps.nval || nflag > ps.nflag || ncoef > ps.ncoef) {
        (void)hipDeviceSynchronize();   // nothing may still use the old buffers
        if (ps.val) (void)hipHostFree(ps.val);
        if (ps.flag) (void)hipHostFree(ps.flag);
        if (ps.coef) (void)hipHostFree(ps.coef);
        ps = PinnedScalars{};
        const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ps.val), sizeof(double) * nval, fl));
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ps.flag), sizeof(int64_t) * nflag, fl));
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ps.coef), sizeof(double) * ncoef, hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&ps.dval), ps.val, 0));
        HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&ps.dflag), ps.flag, 0));
        std::memset(ps.flag, 0, sizeof(int64_t) * nflag);
        ps.nval = nval;
        ps.nflag = nflag;
        ps.ncoef = ncoef;
    }
    return ps;
}

// Waits until flag[r] == seq for every listed run.  The stream is queried
// now and then: if it has drained and a flag is still short, the kernel that
// should have raised it did not, and the wait ends with an error.
void wait_flags(const int64_t *flag, const std::vector<int32_t> &cols, int64_t seq, hipStream_t st)
{
    for (size_t k = 0; k < cols.size();) {
        if (__atomic_load_n(flag + cols[k], __ATOMIC_ACQUIRE) == seq) { ++k; continue; }
        for (int spin = 0; spin < 4096; ++spin)
            if (__atomic_load_n(flag + cols[k], __ATOMIC_ACQUIRE) == seq) break;
        if (__atomic_load_n(flag + cols[k], __ATOMIC_ACQUIRE) == seq) { ++k; continue; }
        const hipError_t q = hipStreamQuery(st);
        if (q == hipErrorNotReady) continue;
        HIPCHK(q);
        if (__atomic_load_n(flag + cols[k], __ATOMIC_ACQUIRE) != seq)
            throw Error("GMRES scalars were not delivered");
    }
}

// ---- R runs of GMRES in lockstep -----------------------------------------
// Slots of the per-run vector table: fixed work vectors, then the Krylov basis.
enum : int { kSlotR = 0, kSlotZ, kSlotWK, kSlotDL, kSlotB, kSlotX, kSlotT1, kSlotT2, kSlotV0 };

// Storage and operators of R runs.  tab[slot*R + r] points at run r's vector
// for the slot (length n[r]); apply() forms out = M^-1 (A in) for the listed
// runs, matvec()/msolve() the two halves.
struct RunSet {
    int R = 0, maxl = 0;
    hipStream_t s = nullptr;           // stream of the batched vector operations
    std::vector<int64_t> n;
    std::vector<int64_t> gap;          // per run: second half's offset (0: contiguous)
    std::vector<double *> htab;
    DBuf<double *> tab;
    DBuf<int64_t> dn;                  // n[0..R), then gap[0..R)
    int nslots() const { return kSlotV0 + maxl + 1; }
    double *ptr(int slot, int r) const { return htab[(size_t)slot * R + r]; }
    void upload_tables()
    {
        tab.upload(htab.data(), htab.size(), s);
        std::vector<int64_t> ng(n);
        gap.resize(n.size(), 0);
        ng.insert(ng.end(), gap.begin(), gap.end());
        dn.upload(ng.data(), ng.size(), s);
    }
    virtual ~RunSet() = default;
    virtual void matvec(const std::vector<int32_t> &cols, const int32_t *dcols, int sin, int sout) = 0;
    virtual void msolve(const std::vector<int32_t> &cols, const int32_t *dcols, int sin, int sout) = 0;
    virtual void apply(const std::vector<int32_t> &cols, const int32_t *dcols, int sin, int sout)
    {
        matvec(cols, dcols, sin, kSlotWK);
        msolve(cols, dcols, kSlotWK, sout);
    }
    // Operators that can take the runs to work on from device flags (one per
    // run, set by device code while a captured graph runs) instead of the host
    // list: returns false where unsupported.  null restores host lists.
    virtual bool device_gate_ok() const { return false; }
    virtual bool use_device_gate(const int32_t *) { return false; }
};

// GMRES with the reference's settings for each run (gmres_dev), the runs in
// lockstep: they share MAXL, so they restart together and differ only in when
// they stop.  Every device operation is batched over the runs still active and
// repeats, per run, the arithmetic of gmres_dev; the host keeps each run's
// Hessenberg matrix, rotations and stopping state.  The B and X slots hold the
// right-hand sides and the initial guesses / solutions.
void gmres_runs(RunSet &rs, DBuf<double> &red, double tol, int nrmax, bool x0_zero, int *iter_out,
                double *err_out, int *ierr_out, bool fallback = false, int *x0_used = nullptr)
{
    const int R = rs.R, maxl = rs.maxl, ldh = maxl + 1;
    hipStream_t st = rs.s;
    PoolScope ps_(st);
    const int64_t cst = maxl + 2;                      // per-run stride of coefficients
    const int64_t dvs = ldh + 4;                       // per-run stride of device scalars
    const int sN = ldh, sS = ldh + 1, sT = ldh + 2;    // slots: |w|^2, |w|^2 after MGS, temp
    int64_t nmax = 0;
    for (int64_t v : rs.n) nmax = std::max(nmax, v);
    const unsigned nb = nblk(nmax, 256);
    DBuf<double> coef;
    coef.alloc((size_t)R * cst);
    red.alloc((size_t)R * dvs * kRedBlocks);          // dot partials per run and slot
    double *const *T = rs.tab.p;
    const int64_t *NV = rs.dn.p;
    auto V = [&](int k) { return kSlotV0 + k; };
    // device copies of run lists
    DBuf<int32_t> dact_a, dact_b;
    std::vector<int32_t> cur_a, cur_b;
    auto put = [&](DBuf<int32_t> &d, std::vector<int32_t> &cur, const std::vector<int32_t> &v) {
        if (v == cur) return;
        cur = v;
        d.upload(v.data(), std::max<size_t>(v.size(), 1), st);
    };
    std::vector<double> hc((size_t)R * cst);
    // Scalars travel through pinned host memory.  The sums the host reads are
    // written there by k_sum_slots, which then raises a per-run flag that the
    // host polls (no copy, no blocking wait); coefficient uploads are
    // asynchronous, each from its own slot of a small ring that is reused only
    // after its copy has completed.
    PinnedScalars &ps = pinned_scalars((size_t)R * dvs, R, kCoefRing * hc.size());
    double *const hv = ps.val;
    int64_t seq = ps.seq;
    EventSet evs(kCoefRing);
    int ring = 0;
    auto sum_fetch = [&](const std::vector<int32_t> &cols, const int32_t *dcols, int s0, int s1, int e0,
                         int e1) {
        ++seq;
        k_sum_slots<<<(unsigned)cols.size(), 256, 0, st>>>(kRedBlocks, red.p, ps.dval, ps.dflag, seq, dvs,
                                                           s0, s1, e0, e1, dcols);
        HIPCHK(hipGetLastError());
        ps.seq = seq;
        wait_flags(ps.flag, cols, seq, st);
    };
    auto put_coef = [&]() {
        double *slot = ps.coef + (size_t)ring * hc.size();
        hipEvent_t done = evs.e[ring];
        spin_wait(done);
        std::memcpy(slot, hc.data(), sizeof(double) * hc.size());
        HIPCHK(hipMemcpyAsync(coef.p, slot, sizeof(double) * hc.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(done, st));
        ring = (ring + 1) % kCoefRing;
    };
    auto gy = [&](size_t cnt) { return dim3(nb, (unsigned)cnt); };
    // hv[r*dvs + slot] = (a, b) for the listed runs, dev_dot's partition and order
    auto dot_fetch = [&](const std::vector<int32_t> &cols, const int32_t *dcols, int sa, int sb, int slot) {
        k_axpy_dott<<<dim3(kRedBlocks, (unsigned)cols.size()), 256, 0, st>>>(T, R, -1, sb, sa, NV, 0,
                                                                             dcols, red.p, dvs, slot, -1);
        sum_fetch(cols, dcols, slot, slot + 1, -1, -1);
    };

    std::vector<int32_t> all(R);
    for (int r = 0; r < R; ++r) all[r] = r;
    put(dact_a, cur_a, all);
    // initial residual
    rs.msolve(all, dact_a.p, kSlotB, kSlotR);
    dot_fetch(all, dact_a.p, kSlotR, kSlotR, sT);
    // M^-1 b is also M^-1 r0 for a run that starts from zero (r0 = b): kept in
    // DL, which nothing writes before the first cycle's residual
    k_copyt<<<gy(R), 256, 0, st>>>(T, R, kSlotR, kSlotDL, NV, dact_a.p);
    std::vector<double> bnrm(R), r0nrm(R), err(R, 0.0), prod(R), rho(R), snormw(R), vnrm(R);
    std::vector<char> guessed(R, 0);        // run started from a nonzero x0
    std::vector<int> iter(R, 0), ierr(R, 0), lgmr(R), iflag(R);
    std::vector<char> conv(R), stall(R), early(R);
    std::vector<std::vector<double>> hes(R, std::vector<double>((size_t)ldh * maxl)),
        q(R, std::vector<double>(2 * maxl));
    for (int r = 0; r < R; ++r) bnrm[r] = std::sqrt(hv[(size_t)r * dvs + sT]);
    {
        std::vector<int32_t> zero, nonzero;
        if (x0_zero) {
            zero = all;
        } else {
            for (int r = 0; r < R; ++r) {
                std::vector<double> hx(rs.n[r]);
                const size_t h = hx.size() / 2;
                HIPCHK(hipMemcpyAsync(hx.data(), rs.ptr(kSlotX, r), sizeof(double) * h,
                                      hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(hx.data() + h, rs.ptr(kSlotX, r) + h + rs.gap[r],
                                      sizeof(double) * (hx.size() - h), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                double sabs = 0.0;
                for (double v : hx) sabs += std::fabs(v);
                (sabs < 1.0e-10 ? zero : nonzero).push_back(r);
            }
        }
        if (!zero.empty()) {
            put(dact_b, cur_b, zero);
            k_copyt<<<gy(zero.size()), 256, 0, st>>>(T, R, kSlotX, kSlotR, NV, dact_b.p);
        }
        if (!nonzero.empty()) {
            put(dact_b, cur_b, nonzero);
            rs.matvec(nonzero, dact_b.p, kSlotX, kSlotR);
        }
        for (int r : nonzero) guessed[r] = 1;
        k_subt<<<gy(R), 256, 0, st>>>(T, R, kSlotB, kSlotR, kSlotR, NV, dact_a.p);
    }
    std::vector<int32_t> active = all;
    int nrsts = 0;
    while (!active.empty()) {
        if (nrsts > nrmax) {
            for (int r : active) ierr[r] = 1;
            break;
        }
        put(dact_a, cur_a, active);
        const unsigned na = (unsigned)active.size();
        if (nrsts > 0) k_copyt<<<gy(na), 256, 0, st>>>(T, R, kSlotDL, kSlotR, NV, dact_a.p);
        k_fillt<<<gy(na), 256, 0, st>>>(T, R, kSlotZ, NV, 0.0, dact_a.p);
        if (nrsts == 0) {
            // M^-1 r0: the kept M^-1 b for runs from zero (a deterministic
            // solve of the same vector), a solve for runs from a guess
            std::vector<int32_t> from0, fromg;
            for (int r : active) (guessed[r] ? fromg : from0).push_back(r);
            if (!from0.empty()) {
                put(dact_b, cur_b, from0);
                k_copyt<<<gy(from0.size()), 256, 0, st>>>(T, R, kSlotDL, kSlotR, NV, dact_b.p);
            }
            if (!fromg.empty()) {
                put(dact_b, cur_b, fromg);
                k_copyt<<<gy(fromg.size()), 256, 0, st>>>(T, R, kSlotR, kSlotWK, NV, dact_b.p);
                rs.msolve(fromg, dact_b.p, kSlotWK, kSlotR);
            }
        }
        k_copyt<<<gy(na), 256, 0, st>>>(T, R, kSlotR, V(0), NV, dact_a.p);
        dot_fetch(active, dact_a.p, V(0), V(0), sT);
        for (int r : active) r0nrm[r] = std::sqrt(hv[(size_t)r * dvs + sT]);
        if (nrsts == 0 && fallback) {
            // an initial guess worse than none: those runs restart from x0 = 0
            std::vector<int32_t> fb;
            for (int r : active)
                if (guessed[r] && !(r0nrm[r] <= bnrm[r])) fb.push_back(r);
            if (!fb.empty()) {
                put(dact_b, cur_b, fb);
                k_fillt<<<gy(fb.size()), 256, 0, st>>>(T, R, kSlotX, NV, 0.0, dact_b.p);
                k_copyt<<<gy(fb.size()), 256, 0, st>>>(T, R, kSlotDL, kSlotR, NV, dact_b.p);  // M^-1 b
                k_copyt<<<gy(fb.size()), 256, 0, st>>>(T, R, kSlotR, V(0), NV, dact_b.p);
                for (int r : fb) { r0nrm[r] = bnrm[r]; guessed[r] = 0; }
            }
        }
        std::vector<int32_t> inner;
        for (int r : active) {
            err[r] = r0nrm[r] / bnrm[r];
            early[r] = err[r] <= tol;
            lgmr[r] = 0;
            iflag[r] = 0;
            conv[r] = 0;
            stall[r] = 0;
            if (!early[r]) inner.push_back(r);
        }
        if (!inner.empty()) {
            put(dact_b, cur_b, inner);
            for (int r : inner) {
                hc[(size_t)r * cst] = 1.0 / r0nrm[r];
                std::fill(hes[r].begin(), hes[r].end(), 0.0);
                prod[r] = 1.0;
                rho[r] = 0.0;
                snormw[r] = 0.0;
            }
            put_coef();
            k_scalt<<<gy(inner.size()), 256, 0, st>>>(T, R, V(0), NV, coef.p, cst, 0, dact_b.p);
        }
        for (int ll = 1; ll <= maxl && !inner.empty(); ++ll) {
            put(dact_b, cur_b, inner);
            const unsigned ni = (unsigned)inner.size();
            for (int r : inner) lgmr[r] = ll;
            rs.apply(inner, dact_b.p, V(ll - 1), V(ll));
            // modified Gram-Schmidt (DORTH): |w|^2 and (V(0), w) in one pass,
            // then each step's w -= h_i V(i) fused with the next step's dot,
            // and the last one with |w|^2
            const dim3 gr(kRedBlocks, ni);
            k_axpy_dott<<<gr, 256, 0, st>>>(T, R, -1, V(ll), V(0), NV, 0, dact_b.p, red.p, dvs, 0, sN);
            for (int i = 1; i <= ll; ++i)
                k_axpy_dott<<<gr, 256, 0, st>>>(T, R, V(i - 1), V(ll), V(i), NV, i - 1, dact_b.p, red.p, dvs,
                                                i < ll ? i : -1, i < ll ? -1 : sS);
            sum_fetch(inner, dact_b.p, 0, ll, sN, sS);
            std::vector<int32_t> reo;
            for (int r : inner) {
                double *hcol = &hes[r][(size_t)(ll - 1) * ldh];
                for (int i = 0; i < ll; ++i) hcol[i] = hv[(size_t)r * dvs + i];
                vnrm[r] = std::sqrt(hv[(size_t)r * dvs + sN]);
                snormw[r] = std::sqrt(hv[(size_t)r * dvs + sS]);
                if (!(vnrm[r] + 0.001 * snormw[r] != vnrm[r])) reo.push_back(r);
            }
            if (!reo.empty()) {
                // conditional re-orthogonalisation (DORTH), one basis vector at a time
                std::vector<double> sumdsq(R, 0.0);
                DBuf<int32_t> dreo, ddr;
                dreo.upload(reo.data(), reo.size(), st);
                for (int i = 0; i < ll; ++i) {
                    std::vector<int32_t> dr;
                    dot_fetch(reo, dreo.p, V(i), V(ll), sT);
                    for (int r : reo) {
                        double *hcol = &hes[r][(size_t)(ll - 1) * ldh];
                        const double tem = -hv[(size_t)r * dvs + sT];
                        if (hcol[i] + 0.001 * tem == hcol[i]) continue;
                        hcol[i] = hcol[i] - tem;
                        hc[(size_t)r * cst + 1] = tem;
                        sumdsq[r] += tem * tem;
                        dr.push_back(r);
                    }
                    if (!dr.empty()) {
                        put_coef();
                        ddr.upload(dr.data(), dr.size(), st);
                        k_axpyt<<<gy(dr.size()), 256, 0, st>>>(T, R, V(i), V(ll), NV, coef.p, cst, 1, 1.0,
                                                              ddr.p);
                    }
                }
                for (int r : reo)
                    if (sumdsq[r] != 0.0)
                        snormw[r] = std::sqrt(std::max(0.0, snormw[r] * snormw[r] - sumdsq[r]));
            }
            std::vector<int32_t> go;
            for (int r : inner) {
                double *hcol = &hes[r][(size_t)(ll - 1) * ldh];
                double *qq = q[r].data();
                hcol[ll] = snormw[r];
                for (int k = 0; k < ll - 1; ++k) {
                    const double cc = qq[2 * k], ss = qq[2 * k + 1];
                    const double t1 = hcol[k], t2 = hcol[k + 1];
                    hcol[k] = cc * t1 - ss * t2;
                    hcol[k + 1] = ss * t1 + cc * t2;
                }
                const double t1 = hcol[ll - 1], t2 = hcol[ll];
                double cc, ss;
                if (t2 == 0.0) { cc = 1.0; ss = 0.0; }
                else if (std::fabs(t2) >= std::fabs(t1)) {
                    const double t = t1 / t2;
                    ss = -1.0 / std::sqrt(1.0 + t * t);
                    cc = -ss * t;
                } else {
                    const double t = t2 / t1;
                    cc = 1.0 / std::sqrt(1.0 + t * t);
                    ss = -cc * t;
                }
                qq[2 * (ll - 1)] = cc;
                qq[2 * (ll - 1) + 1] = ss;
                hcol[ll - 1] = cc * t1 - ss * t2;
                if (hcol[ll - 1] == 0.0) { stall[r] = 1; continue; }
                prod[r] = prod[r] * qq[2 * ll - 1];
                rho[r] = std::fabs(prod[r] * r0nrm[r]);
                // DRLCAL coefficients
                double *cf = &hc[(size_t)r * cst];
                for (int k = 0; k <= ll; ++k) cf[k] = 0.0;
                cf[0] = 1.0;
                for (int i = 1; i <= ll - 1; ++i) {
                    const double s2 = qq[2 * i - 1], c2 = qq[2 * i - 2];
                    for (int k = 0; k < i; ++k) cf[k] *= s2;
                    cf[i] = c2;
                }
                const double s2 = qq[2 * ll - 1], c2 = qq[2 * ll - 2] / snormw[r];
                for (int k = 0; k < ll; ++k) cf[k] *= s2;
                cf[ll] = c2;
                const double tem = r0nrm[r] * prod[r];
                for (int k = 0; k <= ll; ++k) cf[k] *= tem;
                go.push_back(r);
            }
            std::vector<int32_t> next;
            if (!go.empty()) {
                // DRLCAL's residual and its norm, and the new basis vector
                // scaled by 1/|w| for the next step, in one pass
                for (int r : go) hc[(size_t)r * cst + maxl + 1] = 1.0 / snormw[r];
                // one run: its coefficients as kernel arguments, no upload
                const bool by_value = go.size() == 1 && cst <= kCoefArg;
                CoefArg ca{};
                if (by_value) std::copy_n(&hc[(size_t)go[0] * cst], cst, ca.v);
                else put_coef();
                put(dact_b, cur_b, go);
                k_lincomb_dott<<<dim3(kRedBlocks, (unsigned)go.size()), 256, 0, st>>>(T, R, V(0), ll + 1,
                    NV, coef.p, cst, kSlotDL, ll < maxl ? maxl + 1 : -1, dact_b.p, red.p, dvs, sT,
                    by_value, ca);
                sum_fetch(go, dact_b.p, sT, sT + 1, -1, -1);
                for (int r : go) {
                    err[r] = std::sqrt(hv[(size_t)r * dvs + sT]) / bnrm[r];
                    if (err[r] <= tol) conv[r] = 1;
                    else next.push_back(r);
                }
            }
            if (ll == maxl) break;
            inner = next;
        }
        // finish the cycle for every active run
        std::vector<int32_t> add_z, still;
        std::map<int, std::vector<int32_t>> by_nn;
        for (int r : active) {
            if (early[r]) { iflag[r] = 0; continue; }
            if (stall[r]) { iflag[r] = 2; continue; }
            if (!conv[r]) iflag[r] = (rho[r] < r0nrm[r]) ? 1 : 2;
            if (iflag[r] != 2) {
                const int nn = lgmr[r];
                std::vector<double> yls(maxl + 1, 0.0);
                yls[0] = r0nrm[r];
                const double *qq = q[r].data();
                for (int k = 0; k < nn; ++k) {
                    const double cc = qq[2 * k], ss = qq[2 * k + 1];
                    const double t1 = yls[k], t2 = yls[k + 1];
                    yls[k] = cc * t1 - ss * t2;
                    yls[k + 1] = ss * t1 + cc * t2;
                }
                const std::vector<double> &H = hes[r];
                for (int kb = 0; kb < nn; ++kb) {
                    const int k = nn - 1 - kb;
                    yls[k] = yls[k] / H[(size_t)k * ldh + k];
                    const double t = -yls[k];
                    for (int i = 0; i < k; ++i) yls[i] += t * H[(size_t)k * ldh + i];
                }
                for (int k = 0; k < nn; ++k) hc[(size_t)r * cst + k] = yls[k];
                by_nn[nn].push_back(r);
            }
        }
        if (!by_nn.empty()) {
            put_coef();
            for (auto &kv : by_nn) {
                DBuf<int32_t> dg;
                dg.upload(kv.second.data(), kv.second.size(), st);
                k_lincombt<<<gy(kv.second.size()), 256, 0, st>>>(T, R, V(0), kv.first, NV, coef.p, cst,
                                                                 kSlotZ, dg.p);
            }
        }
        for (int r : active) {
            iter[r] += lgmr[r];
            if (iflag[r] != 2) add_z.push_back(r);
        }
        if (!add_z.empty()) {
            for (int r : add_z) hc[(size_t)r * cst + maxl + 1] = 1.0;
            put_coef();
            put(dact_b, cur_b, add_z);
            k_axpyt<<<gy(add_z.size()), 256, 0, st>>>(T, R, kSlotZ, kSlotX, NV, coef.p, cst, maxl + 1, 1.0,
                                                     dact_b.p);
        }
        for (int r : active) {
            if (iflag[r] == 0) ierr[r] = 0;
            else if (iflag[r] == 1) still.push_back(r);
            else ierr[r] = 2;
        }
        active = still;
        ++nrsts;
    }
    HIPCHK(hipStreamSynchronize(st));
    for (int r = 0; r < R; ++r) {
        iter_out[r] = iter[r];
        err_out[r] = err[r];
        ierr_out[r] = ierr[r];
        if (x0_used) x0_used[r] = guessed[r];
    }
}

// ---- GMRES with its control on the device (one captured graph per cycle) ----
// The scalar work of gmres_runs -- the Hessenberg column, the conditional
// re-orthogonalisation, the Givens rotations, DRLCAL's coefficients, the
// stopping test, DHELS and the restart decision -- done by small device
// kernels on per-run state in device memory, with exactly the host's
// operations (no contraction, the same summation of the dot partials), so a
// whole restart cycle (MAXL Arnoldi steps and the solution update) is one
// captured graph that is launched once per cycle for every run at once.  The
// runs each kernel works on come from device-built lists with device counts;
// the operator skips the work of runs no longer iterating through device gate
// flags; the host only waits for the count of runs still active after each
// cycle.  Results are bit-identical to gmres_runs.
constexpr int kGmMaxl = 32;    // largest MAXL of the device form
enum : int { kLstC = 0, kLstA, kLstR, kLstG, kLstY, kLstN };
enum : int { kFlA = 1, kFlReo = 2, kFlConv = 4, kFlStall = 8, kFlEarly = 16 };

struct GmDev {
    int R, maxl, ldh, cst, dvs, np, nrmax;
    double tol;
    double *bnrm, *r0nrm, *err, *prod, *rho, *snormw, *vnrm, *hes, *q, *coef;
    const double *part;
    int32_t *iter, *ierr, *lgmr, *iflag, *fl, *act, *gate, *nvv, *lst, *cnt;
    double *hval;          // mapped host memory: number of runs still active
    int64_t *hflag;        // mapped host memory: raised after each cycle
    __device__ double *H(int r) const { return hes + (size_t)r * ldh * maxl; }
    __device__ double *Q(int r) const { return q + (size_t)r * 2 * maxl; }
    __device__ double *CF(int r) const { return coef + (size_t)r * cst; }
    __device__ const double *P(int r, int slot) const { return part + ((size_t)r * dvs + slot) * np; }
};

// lists[kind] = the runs of lists[from] (in order) whose flags have bit `bit`,
// or (from < 0) the runs r = 0..R-1 with act[r]; optionally gate[r] = membership
__global__ void k_gm_list(GmDev g, int kind, int from, int bit, int set_gate)
{
    if (threadIdx.x != 0) return;
    int *out = g.lst + (size_t)kind * g.R, n = 0;
    if (set_gate)
        for (int r = 0; r < g.R; ++r) g.gate[r] = 0;
    const int cnt = from < 0 ? g.R : g.cnt[from];
    for (int k = 0; k < cnt; ++k) {
        const int r = from < 0 ? k : g.lst[(size_t)from * g.R + k];
        const bool in = from < 0 ? g.act[r] != 0 : (g.fl[r] & bit) != 0;
        if (!in) continue;
        out[n++] = r;
        if (set_gate) g.gate[r] = 1;
    }
    g.cnt[kind] = n;
}

// cycle start, per run of C (gm.f:1652-1700): r0nrm = |M^-1 r0|, early exit,
// the first basis vector's scale, cleared Hessenberg matrix
__global__ void __launch_bounds__(256) k_gm_start(GmDev g, int sT)
{
#pragma clang fp contract(off)
    __shared__ double red[256];
    if ((int)blockIdx.x >= g.cnt[kLstC]) return;
    const int r = g.lst[(size_t)kLstC * g.R + blockIdx.x];
    const double s = sum_partials(g.P(r, sT), g.np, red);
    const double rn = sqrt(s);
    const bool early = rn / g.bnrm[r] <= g.tol;
    if (!early)
        for (int k = threadIdx.x; k < g.ldh * g.maxl; k += blockDim.x) g.H(r)[k] = 0.0;
    if (threadIdx.x != 0) return;
    g.r0nrm[r] = rn;
    g.err[r] = rn / g.bnrm[r];
    g.lgmr[r] = 0;
    g.iflag[r] = 0;
    g.fl[r] = early ? kFlEarly : kFlA;
    if (!early) {
        g.CF(r)[0] = 1.0 / rn;
        g.prod[r] = 1.0;
        g.rho[r] = 0.0;
        g.snormw[r] = 0.0;
    }
}

// Arnoldi step ll, per run of A: the Hessenberg column from the modified
// Gram-Schmidt dots, |w| before and after, and DORTH's test for the
// re-orthogonalisation pass
__global__ void __launch_bounds__(256) k_gm_hcol(GmDev g, int ll, int sN, int sS)
{
#pragma clang fp contract(off)
    __shared__ double red[256];
    if ((int)blockIdx.x >= g.cnt[kLstA]) return;
    const int r = g.lst[(size_t)kLstA * g.R + blockIdx.x];
    double *hcol = g.H(r) + (size_t)(ll - 1) * g.ldh;
    for (int i = 0; i < ll; ++i) {
        const double v = sum_partials(g.P(r, i), g.np, red);
        if (threadIdx.x == 0) hcol[i] = v;
    }
    const double vn = sqrt(sum_partials(g.P(r, sN), g.np, red));
    const double sw = sqrt(sum_partials(g.P(r, sS), g.np, red));
    if (threadIdx.x != 0) return;
    g.lgmr[r] = ll;
    g.vnrm[r] = vn;
    g.snormw[r] = sw;
    if (!(vn + 0.001 * sw != vn)) g.fl[r] |= kFlReo;
}

// after the re-orthogonalisation chain, per run of the list R: DORTH's updates
// of the column (TEM of each basis vector from its dot) and of |w|
__global__ void __launch_bounds__(256) k_gm_reo_end(GmDev g, int ll)
{
#pragma clang fp contract(off)
    __shared__ double red[256];
    if ((int)blockIdx.x >= g.cnt[kLstR]) return;
    const int r = g.lst[(size_t)kLstR * g.R + blockIdx.x];
    double *hcol = g.H(r) + (size_t)(ll - 1) * g.ldh;
    double sumdsq = 0.0;
    for (int i = 0; i < ll; ++i) {
        const double tem = -sum_partials(g.P(r, i), g.np, red);
        if (threadIdx.x == 0 && !(hcol[i] + 0.001 * tem == hcol[i])) {
            hcol[i] = hcol[i] - tem;
            sumdsq += tem * tem;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (sumdsq != 0.0) {
        const double sw = g.snormw[r];
        g.snormw[r] = sqrt(fmax(0.0, sw * sw - sumdsq));
    }
    g.fl[r] &= ~kFlReo;
}

// Givens rotations of the new column (DHEQR, gm.f:1093-1125), the stall test,
// the residual estimate and DRLCAL's coefficients (gm.f:1948-1977), one thread
// per run of A; then the list G of runs that form DRLCAL's residual
__global__ void __launch_bounds__(256) k_gm_givens(GmDev g, int ll)
{
#pragma clang fp contract(off)
    const int na = g.cnt[kLstA];
    for (int b = threadIdx.x; b < na; b += blockDim.x) {
        const int r = g.lst[(size_t)kLstA * g.R + b];
        double *hcol = g.H(r) + (size_t)(ll - 1) * g.ldh;
        double *qq = g.Q(r);
        hcol[ll] = g.snormw[r];
        for (int k = 0; k < ll - 1; ++k) {
            const double cc = qq[2 * k], ss = qq[2 * k + 1];
            const double t1 = hcol[k], t2 = hcol[k + 1];
            hcol[k] = cc * t1 - ss * t2;
            hcol[k + 1] = ss * t1 + cc * t2;
        }
        const double t1 = hcol[ll - 1], t2 = hcol[ll];
        double cc, ss;
        if (t2 == 0.0) { cc = 1.0; ss = 0.0; }
        else if (fabs(t2) >= fabs(t1)) {
            const double t = t1 / t2;
            ss = -1.0 / sqrt(1.0 + t * t);
            cc = -ss * t;
        } else {
            const double t = t2 / t1;
            cc = 1.0 / sqrt(1.0 + t * t);
            ss = -cc * t;
        }
        qq[2 * (ll - 1)] = cc;
        qq[2 * (ll - 1) + 1] = ss;
        hcol[ll - 1] = cc * t1 - ss * t2;
        if (hcol[ll - 1] == 0.0) { g.fl[r] = (g.fl[r] & ~kFlA) | kFlStall; continue; }
        g.prod[r] = g.prod[r] * qq[2 * ll - 1];
        g.rho[r] = fabs(g.prod[r] * g.r0nrm[r]);
        double *cf = g.CF(r);
        for (int k = 0; k <= ll; ++k) cf[k] = 0.0;
        cf[0] = 1.0;
        for (int i = 1; i <= ll - 1; ++i) {
            const double s2 = qq[2 * i - 1], c2 = qq[2 * i - 2];
            for (int k = 0; k < i; ++k) cf[k] *= s2;
            cf[i] = c2;
        }
        const double s2 = qq[2 * ll - 1], c2 = qq[2 * ll - 2] / g.snormw[r];
        for (int k = 0; k < ll; ++k) cf[k] *= s2;
        cf[ll] = c2;
        const double tem = g.r0nrm[r] * g.prod[r];
        for (int k = 0; k <= ll; ++k) cf[k] *= tem;
        cf[g.maxl + 1] = 1.0 / g.snormw[r];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int n = 0;
    int *out = g.lst + (size_t)kLstG * g.R;
    for (int b = 0; b < na; ++b) {
        const int r = g.lst[(size_t)kLstA * g.R + b];
        if (!(g.fl[r] & kFlStall)) out[n++] = r;
    }
    g.cnt[kLstG] = n;
}

// DRLCAL's residual norm and the stopping test (ITOL=2), per run of G
__global__ void __launch_bounds__(256) k_gm_conv(GmDev g, int sT)
{
#pragma clang fp contract(off)
    __shared__ double red[256];
    if ((int)blockIdx.x >= g.cnt[kLstG]) return;
    const int r = g.lst[(size_t)kLstG * g.R + blockIdx.x];
    const double s = sum_partials(g.P(r, sT), g.np, red);
    if (threadIdx.x != 0) return;
    g.err[r] = sqrt(s) / g.bnrm[r];
    if (g.err[r] <= g.tol) g.fl[r] = (g.fl[r] & ~kFlA) | kFlConv;
}

// end of a cycle, per run of C: the outcome (DPIGMR's IFLAG), DHELS's least
// squares solution (gm.f:925-947) as the coefficients of the update z =
// V y; then the list Y of runs to update
__global__ void __launch_bounds__(256) k_gm_end(GmDev g)
{
#pragma clang fp contract(off)
    const int nc = g.cnt[kLstC];
    for (int b = threadIdx.x; b < nc; b += blockDim.x) {
        const int r = g.lst[(size_t)kLstC * g.R + b];
        const int f = g.fl[r];
        int iflag;
        if (f & kFlEarly) iflag = 0;
        else if (f & kFlStall) iflag = 2;
        else if (f & kFlConv) iflag = 0;
        else iflag = (g.rho[r] < g.r0nrm[r]) ? 1 : 2;
        g.iflag[r] = iflag;
        g.nvv[r] = 0;
        if ((f & kFlEarly) || iflag == 2) continue;
        const int nn = g.lgmr[r];
        double yls[kGmMaxl + 1];
        for (int k = 0; k <= g.maxl; ++k) yls[k] = 0.0;
        yls[0] = g.r0nrm[r];
        const double *qq = g.Q(r);
        for (int k = 0; k < nn; ++k) {
            const double cc = qq[2 * k], ss = qq[2 * k + 1];
            const double t1 = yls[k], t2 = yls[k + 1];
            yls[k] = cc * t1 - ss * t2;
            yls[k + 1] = ss * t1 + cc * t2;
        }
        const double *Hm = g.H(r);
        for (int kb = 0; kb < nn; ++kb) {
            const int k = nn - 1 - kb;
            yls[k] = yls[k] / Hm[(size_t)k * g.ldh + k];
            const double t = -yls[k];
            for (int i = 0; i < k; ++i) yls[i] += t * Hm[(size_t)k * g.ldh + i];
        }
        double *cf = g.CF(r);
        for (int k = 0; k < nn; ++k) cf[k] = yls[k];
        cf[g.maxl + 1] = 1.0;
        g.nvv[r] = nn;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int n = 0;
    int *out = g.lst + (size_t)kLstY * g.R;
    for (int b = 0; b < nc; ++b) {
        const int r = g.lst[(size_t)kLstC * g.R + b];
        if (g.nvv[r] > 0) out[n++] = r;
    }
    g.cnt[kLstY] = n;
}

// after the update x += z: iteration counts, IERR, the runs that restart, the
// NRMAX cap (IERR=1, gm.f:821-850); the count of active runs goes to the host
__global__ void k_gm_fin(GmDev g, int64_t seq)
{
    if (threadIdx.x != 0) return;
    const int nc = g.cnt[kLstC];
    int *nrsts = g.cnt + kLstN;
    for (int b = 0; b < nc; ++b) {
        const int r = g.lst[(size_t)kLstC * g.R + b];
        g.iter[r] += g.lgmr[r];
        if (g.iflag[r] == 1) continue;
        g.ierr[r] = g.iflag[r] == 0 ? 0 : 2;
        g.act[r] = 0;
    }
    *nrsts += 1;
    int na = 0;
    for (int r = 0; r < g.R; ++r) {
        if (!g.act[r]) continue;
        if (*nrsts > g.nrmax) { g.ierr[r] = 1; g.act[r] = 0; continue; }
        ++na;
    }
    g.hval[0] = (double)na;
    __threadfence_system();
    __hip_atomic_store(g.hflag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The runs of rs from their initial residual state (DL = M^-1 r0, X = x0,
// bnrm = |M^-1 b|, iter/ierr zero) to the end, cycle by cycle through one
// captured graph; returns false (nothing done) when rs cannot take device gates.
bool gmres_graph_cycles(RunSet &rs, DBuf<double> &red, double tol, int nrmax,
                        const std::vector<double> &bnrm, int *iter_out, double *err_out,
                        int *ierr_out, int64_t *launches)
{
    const int R = rs.R, maxl = rs.maxl, ldh = maxl + 1;
    if (maxl > kGmMaxl || maxl < 1) return false;
    hipStream_t st = rs.s;
    const int64_t cst = maxl + 2, dvs = ldh + 4;
    const int sN = ldh, sS = ldh + 1, sT = ldh + 2;
    int64_t nmax = 0;
    for (int64_t v : rs.n) nmax = std::max(nmax, v);
    const unsigned nb = nblk(nmax, 256);
    DBuf<double> dd;
    DBuf<int32_t> di;
    const size_t nd = (size_t)R * (7 + ldh * maxl + 2 * maxl + cst);
    const size_t ni = (size_t)R * (8 + kLstN) + kLstN + 1;
    dd.alloc(nd);
    di.alloc(ni);
    red.alloc((size_t)R * dvs * kRedBlocks);
    GmDev g{};
    g.R = R; g.maxl = maxl; g.ldh = ldh; g.cst = (int)cst; g.dvs = (int)dvs; g.np = kRedBlocks;
    g.nrmax = nrmax; g.tol = tol;
    double *p = dd.p;
    g.bnrm = p; p += R; g.r0nrm = p; p += R; g.err = p; p += R; g.prod = p; p += R;
    g.rho = p; p += R; g.snormw = p; p += R; g.vnrm = p; p += R;
    g.hes = p; p += (size_t)R * ldh * maxl; g.q = p; p += (size_t)R * 2 * maxl; g.coef = p;
    int32_t *ip = di.p;
    g.iter = ip; ip += R; g.ierr = ip; ip += R; g.lgmr = ip; ip += R; g.iflag = ip; ip += R;
    g.fl = ip; ip += R; g.act = ip; ip += R; g.gate = ip; ip += R; g.nvv = ip; ip += R;
    g.lst = ip; ip += (size_t)R * kLstN; g.cnt = ip;
    g.part = red.p;
    // initial state: every run active, counters zero
    {
        std::vector<double> hb(bnrm);
        std::vector<int32_t> hi(ni, 0);
        for (int r = 0; r < R; ++r) hi[(size_t)5 * R + r] = 1;   // act
        HIPCHK(hipMemcpyAsync(g.bnrm, hb.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(di.p, hi.data(), sizeof(int32_t) * ni, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    PinnedScalars &ps = pinned_scalars(std::max<size_t>(1, (size_t)R * dvs), std::max(R, 1), (size_t)R * cst * kCoefRing);
    g.hval = ps.dval;
    g.hflag = ps.dflag;
    if (!rs.use_device_gate(g.gate)) return false;
    double *const *T = rs.tab.p;
    const int64_t *NV = rs.dn.p;
    auto V = [&](int k) { return kSlotV0 + k; };
    const int32_t *LC = g.lst + (size_t)kLstC * R, *LA = g.lst + (size_t)kLstA * R,
                  *LR = g.lst + (size_t)kLstR * R, *LG = g.lst + (size_t)kLstG * R,
                  *LY = g.lst + (size_t)kLstY * R;
    const int32_t *nC = g.cnt + kLstC, *nA = g.cnt + kLstA, *nR = g.cnt + kLstR,
                  *nG = g.cnt + kLstG, *nY = g.cnt + kLstY;
    const dim3 gv(nb, (unsigned)R), gr(kRedBlocks, (unsigned)R);
    const std::vector<int32_t> none;
    int64_t seq0 = ps.seq;
    // one cycle of launches (captured once, replayed per cycle); the flag
    // value it raises is read from a device counter-free sequence: each replay
    // raises seq0 + 1 and the host resets the flag before the next
    auto cycle = [&]() {
        k_gm_list<<<1, 64, 0, st>>>(g, kLstC, -1, 0, 0);
        k_copyt<<<gv, 256, 0, st>>>(T, R, kSlotDL, kSlotR, NV, LC, nC);
        k_fillt<<<gv, 256, 0, st>>>(T, R, kSlotZ, NV, 0.0, LC, nC);
        k_copyt<<<gv, 256, 0, st>>>(T, R, kSlotR, V(0), NV, LC, nC);
        k_axpy_dott<<<gr, 256, 0, st>>>(T, R, -1, V(0), V(0), NV, 0, LC, red.p, dvs, sT, -1, nC);
        k_gm_start<<<R, 256, 0, st>>>(g, sT);
        k_gm_list<<<1, 64, 0, st>>>(g, kLstA, kLstC, kFlA, 1);
        k_scalt<<<gv, 256, 0, st>>>(T, R, V(0), NV, g.coef, cst, 0, LA, nA);
        for (int ll = 1; ll <= maxl; ++ll) {
            rs.apply(none, LA, V(ll - 1), V(ll));
            k_axpy_dott<<<gr, 256, 0, st>>>(T, R, -1, V(ll), V(0), NV, 0, LA, red.p, dvs, 0, sN, nA);
            for (int i = 1; i <= ll; ++i)
                k_axpy_dott<<<gr, 256, 0, st>>>(T, R, V(i - 1), V(ll), V(i), NV, i - 1, LA, red.p, dvs,
                                                i < ll ? i : -1, i < ll ? -1 : sS, nA);
            k_gm_hcol<<<R, 256, 0, st>>>(g, ll, sN, sS);
            // DORTH's second pass for the runs that need it (rare): dot i,
            // then the conditional axpy of i fused with dot i + 1
            k_gm_list<<<1, 64, 0, st>>>(g, kLstR, kLstA, kFlReo, 0);
            k_axpy_dott<<<gr, 256, 0, st>>>(T, R, -1, V(ll), V(0), NV, 0, LR, red.p, dvs, 0, -1, nR);
            for (int i = 1; i <= ll; ++i)
                k_axpy_dott<<<gr, 256, 0, st>>>(T, R, V(i - 1), V(ll), V(i), NV, i - 1, LR, red.p, dvs,
                                                i < ll ? i : -1, -1, nR,
                                                g.hes + (size_t)(ll - 1) * ldh + (i - 1), (int64_t)ldh * maxl);
            k_gm_reo_end<<<R, 256, 0, st>>>(g, ll);
            k_gm_givens<<<1, 256, 0, st>>>(g, ll);
            k_lincomb_dott<<<gr, 256, 0, st>>>(T, R, V(0), ll + 1, NV, g.coef, cst, kSlotDL,
                                               ll < maxl ? maxl + 1 : -1, LG, red.p, dvs, sT, false,
                                               CoefArg{}, nG);
            k_gm_conv<<<R, 256, 0, st>>>(g, sT);
            k_gm_list<<<1, 64, 0, st>>>(g, kLstA, kLstA, kFlA, 1);
        }
        k_gm_end<<<1, 256, 0, st>>>(g);
        k_lincombt<<<gv, 256, 0, st>>>(T, R, V(0), 0, NV, g.coef, cst, kSlotZ, LY, nY, g.nvv);
        k_axpyt<<<gv, 256, 0, st>>>(T, R, kSlotZ, kSlotX, NV, g.coef, cst, maxl + 1, 1.0, LY, nY);
        k_gm_fin<<<1, 64, 0, st>>>(g, seq0 + 1);
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        cycle();
    } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        rs.use_device_gate(nullptr);
        throw;
    }
    HIPCHK(hipStreamEndCapture(st, &graph));
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        rs.use_device_gate(nullptr);
        HIPCHK(ie);
    }
    int64_t nl = 0;
    try {
        for (;;) {
            __atomic_store_n(ps.flag, seq0, __ATOMIC_RELEASE);
            HIPCHK(hipGraphLaunch(exec, st));
            ++nl;
            wait_flags(ps.flag, std::vector<int32_t>{0}, seq0 + 1, st);
            if (ps.val[0] == 0.0) break;
        }
    } catch (...) {
        (void)hipStreamSynchronize(st);
        ps.seq = seq0 + 1;
        (void)hipGraphExecDestroy(exec);
        rs.use_device_gate(nullptr);
        throw;
    }
    HIPCHK(hipStreamSynchronize(st));
    ps.seq = seq0 + 1;                   // the flag now holds this value
    (void)hipGraphExecDestroy(exec);
    rs.use_device_gate(nullptr);
    std::vector<int32_t> it(R), ier(R);
    HIPCHK(hipMemcpy(it.data(), g.iter, sizeof(int32_t) * R, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ier.data(), g.ierr, sizeof(int32_t) * R, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(err_out, g.err, sizeof(double) * R, hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r) { iter_out[r] = it[r]; ierr_out[r] = ier[r]; }
    if (launches) *launches = nl;
    return true;
}

// gmres_runs with the cycles on the device (gmres_graph_cycles): the host
// forms each run's initial state as gmres_runs does -- M^-1 b and its norm,
// the initial guess's preconditioned residual M^-1 (b - A x0), the fallback
// to x0 = 0 when that residual is the larger -- and leaves M^-1 r0 in DL.
// Returns false (nothing done) for run sets without device gates or MAXL
// above kGmMaxl.
bool gmres_graph(RunSet &rs, DBuf<double> &red, double tol, int nrmax, bool x0_zero, int *iter_out,
                 double *err_out, int *ierr_out, bool fallback = false, int *x0_used = nullptr,
                 int64_t *launches = nullptr)
{
    const int R = rs.R, maxl = rs.maxl, ldh = maxl + 1;
    if (!rs.device_gate_ok() || maxl > kGmMaxl || maxl < 1) return false;
    hipStream_t st = rs.s;
    PoolScope ps_(st);
    const int64_t dvs = ldh + 4;
    const int sT = ldh + 2;
    int64_t nmax = 0;
    for (int64_t v : rs.n) nmax = std::max(nmax, v);
    const unsigned nb = nblk(nmax, 256);
    red.alloc((size_t)R * dvs * kRedBlocks);
    double *const *T = rs.tab.p;
    const int64_t *NV = rs.dn.p;
    PinnedScalars &ps = pinned_scalars((size_t)R * dvs, R, (size_t)kCoefRing * R * (maxl + 2));
    auto gy = [&](size_t cnt) { return dim3(nb, (unsigned)cnt); };
    auto dots = [&](const std::vector<int32_t> &cols, const int32_t *dcols, int sa, int sb) {
        k_axpy_dott<<<dim3(kRedBlocks, (unsigned)cols.size()), 256, 0, st>>>(T, R, -1, sb, sa, NV, 0,
                                                                             dcols, red.p, dvs, sT, -1);
        ++ps.seq;
        k_sum_slots<<<(unsigned)cols.size(), 256, 0, st>>>(kRedBlocks, red.p, ps.dval, ps.dflag, ps.seq,
                                                            dvs, sT, sT + 1, -1, -1, dcols);
        HIPCHK(hipGetLastError());
        wait_flags(ps.flag, cols, ps.seq, st);
    };
    std::vector<int32_t> all(R);
    for (int r = 0; r < R; ++r) all[r] = r;
    DBuf<int32_t> dall, dsub;
    dall.upload(all.data(), R, st);
    rs.msolve(all, dall.p, kSlotB, kSlotR);
    dots(all, dall.p, kSlotR, kSlotR);
    std::vector<double> bnrm(R);
    for (int r = 0; r < R; ++r) bnrm[r] = std::sqrt(ps.val[(size_t)r * dvs + sT]);
    k_copyt<<<gy(R), 256, 0, st>>>(T, R, kSlotR, kSlotDL, NV, dall.p);
    std::vector<int32_t> guessed;
    if (!x0_zero) {
        for (int r = 0; r < R; ++r) {
            std::vector<double> hx(rs.n[r]);
            const size_t h = hx.size() / 2;
            HIPCHK(hipMemcpyAsync(hx.data(), rs.ptr(kSlotX, r), sizeof(double) * h, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(hx.data() + h, rs.ptr(kSlotX, r) + h + rs.gap[r],
                                  sizeof(double) * (hx.size() - h), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            double sabs = 0.0;
            for (double v : hx) sabs += std::fabs(v);
            if (!(sabs < 1.0e-10)) guessed.push_back(r);
        }
    }
    std::vector<int> used(R, 0);
    if (!guessed.empty()) {
        // M^-1 (b - A x0) into Z, kept in DL unless it is worse than M^-1 b
        dsub.upload(guessed.data(), guessed.size(), st);
        rs.matvec(guessed, dsub.p, kSlotX, kSlotR);
        k_subt<<<gy(guessed.size()), 256, 0, st>>>(T, R, kSlotB, kSlotR, kSlotR, NV, dsub.p);
        k_copyt<<<gy(guessed.size()), 256, 0, st>>>(T, R, kSlotR, kSlotWK, NV, dsub.p);
        rs.msolve(guessed, dsub.p, kSlotWK, kSlotZ);
        dots(guessed, dsub.p, kSlotZ, kSlotZ);
        std::vector<int32_t> keep, drop;
        for (int r : guessed) {
            const double r0 = std::sqrt(ps.val[(size_t)r * dvs + sT]);
            (fallback && !(r0 <= bnrm[r]) ? drop : keep).push_back(r);
        }
        if (!keep.empty()) {
            DBuf<int32_t> dk;
            dk.upload(keep.data(), keep.size(), st);
            k_copyt<<<gy(keep.size()), 256, 0, st>>>(T, R, kSlotZ, kSlotDL, NV, dk.p);
            for (int r : keep) used[r] = 1;
            HIPCHK(hipStreamSynchronize(st));
        }
        if (!drop.empty()) {
            DBuf<int32_t> dd;
            dd.upload(drop.data(), drop.size(), st);
            k_fillt<<<gy(drop.size()), 256, 0, st>>>(T, R, kSlotX, NV, 0.0, dd.p);
            HIPCHK(hipStreamSynchronize(st));
        }
    }
    HIPCHK(hipGetLastError());
    if (!gmres_graph_cycles(rs, red, tol, nrmax, bnrm, iter_out, err_out, ierr_out, launches))
        throw Error("device GMRES unavailable for this run set");
    if (x0_used)
        for (int r = 0; r < R; ++r) x0_used[r] = used[r];
    return true;
}

// R right-hand sides on one surface: vectors stored with a common stride so the
// block field and preconditioner kernels see them as block columns
struct BlockRuns : RunSet {
    tabi_ctx *c;
    int64_t N;
    DBuf<double> V, work;
    std::vector<int64_t> stride;       // per slot: distance between runs
    BlockRuns(tabi_ctx *c_, int R_, int maxl_, double *dB, double *dX) : c(c_)
    {
        R = R_;
        maxl = maxl_;
        s = c->s;
        N = 2 * (int64_t)c->n;
        n.assign(R, N);
        const int64_t vst = (int64_t)(maxl + 1) * N;
        V.alloc((size_t)R * vst);
        work.alloc((size_t)R * N * 4);
        htab.assign((size_t)nslots() * R, nullptr);
        stride.assign(nslots(), N);
        for (int r = 0; r < R; ++r) {
            htab[(size_t)kSlotR * R + r] = work.p + (size_t)(0 * R + r) * N;
            htab[(size_t)kSlotZ * R + r] = work.p + (size_t)(1 * R + r) * N;
            htab[(size_t)kSlotWK * R + r] = work.p + (size_t)(2 * R + r) * N;
            htab[(size_t)kSlotDL * R + r] = work.p + (size_t)(3 * R + r) * N;
            htab[(size_t)kSlotB * R + r] = dB + (size_t)r * N;
            htab[(size_t)kSlotX * R + r] = dX + (size_t)r * N;
            for (int k = 0; k <= maxl; ++k)
                htab[(size_t)(kSlotV0 + k) * R + r] = V.p + (size_t)r * vst + (size_t)k * N;
        }
        for (int k = 0; k <= maxl; ++k) stride[kSlotV0 + k] = vst;
        upload_tables();
    }
    void matvec(const std::vector<int32_t> &cols, const int32_t *dcols, int sin, int sout) override
    {
        matvec_block_dev(c, (int)cols.size(), ptr(sin, 0), stride[sin], ptr(sout, 0), stride[sout], dcols);
    }
    void msolve(const std::vector<int32_t> &cols, const int32_t *dcols, int sin, int sout) override
    {
        msolve_block_dev(c, (int)cols.size(), ptr(sin, 0), stride[sin], ptr(sout, 0), stride[sout], dcols);
    }
};

void gmres_block_dev(tabi_ctx *c, int R, double *dB, double *dX, double tol, int maxl, int nrmax,
                     bool x0_zero, int *iter_out, double *err_out, int *ierr_out)
{
    const int64_t N = 2 * (int64_t)c->n;
    if (maxl > N) maxl = (int)N;
    BlockRuns rs(c, R, maxl, dB, dX);
    gmres_runs(rs, c->blk_red, tol, nrmax, x0_zero, iter_out, err_out, ierr_out);
}

// Right-hand side of the faces [f0, f0 + nf) (default: all of the context's):
// b[f0 + i] and b[ho + f0 + i], ho = the context's n (a packed context's
// vectors hold every surface's first halves, then every second half).
void rhs_dev(tabi_ctx *c, int nq, const double *qpos, const double *q, double eps_p, double *db,
             bool dev = false, int f0 = 0, int nf = -1, bool sync = true)
{
    if (nf < 0) nf = c->n;
    DevIn dqp(qpos, (size_t)3 * nq, dev, c->s), dq(q, nq, dev, c->s);
    if (c->prm.precision == 1) {
        DBuf<QRec32> qrec;
        qrec.alloc(std::max(nq, 1));
        if (nq > 0) k_charge_rec32<<<nblk(nq, 256), 256, 0, c->s>>>(nq, dqp.p, dq.p, qrec.p);
        k_rhs32<<<nblk(nf, 256), 256, 0, c->s>>>(nf, nq, qrec.p, eps_p, c->tx.p + f0, c->ty.p + f0,
            c->tz.p + f0, c->tnx.p + f0, c->tny.p + f0, c->tnz.p + f0, db + f0, c->n);
        HIPCHK(hipGetLastError());
        if (sync) HIPCHK(hipStreamSynchronize(c->s));
        return;
    }
    DBuf<double4> qrec;
    qrec.alloc(std::max(nq, 1));
    if (nq > 0) k_charge_rec<<<nblk(nq, 256), 256, 0, c->s>>>(nq, dqp.p, dq.p, qrec.p);
    k_rhs<<<nblk(nf, 256), 256, 0, c->s>>>(nf, nq, qrec.p, eps_p, c->tx.p + f0, c->ty.p + f0,
        c->tz.p + f0, c->tnx.p + f0, c->tny.p + f0, c->tnz.p + f0, db + f0, c->n);
    HIPCHK(hipGetLastError());
    if (sync) HIPCHK(hipStreamSynchronize(c->s));
}

// Many surfaces, one context each (own tree, lists and factors), one GMRES run
// per surface, all in lockstep: the batched vector operations run on a common
// stream, and each step's matvec + preconditioner solve of every active surface
// is replayed from one captured graph whose branches (one per surface, on the
// surface's own stream) run concurrently, so many small surfaces fill the GPU
// together.  The graph is captured again when the set of active surfaces changes.
struct MultiRuns : RunSet {
    std::vector<tabi_ctx *> cs;
    std::vector<std::unique_ptr<DBuf<double>>> store;
    bool use_graph = true;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::map<std::vector<int32_t>, hipGraphExec_t> graphs;
    int64_t graph_launches = 0, captures = 0;

    MultiRuns(std::vector<tabi_ctx *> cs_, int maxl_, hipStream_t main, bool graph,
              const std::vector<double *> &dB, const std::vector<double *> &dX)
        : cs(std::move(cs_)), use_graph(graph)
    {
        R = (int)cs.size();
        s = main;
        maxl = maxl_;
        for (auto *c : cs) maxl = std::min<int64_t>(maxl, 2 * (int64_t)c->n);
        n.resize(R);
        htab.assign((size_t)nslots() * R, nullptr);
        for (int r = 0; r < R; ++r) {
            tabi_ctx *c = cs[r];
            const int64_t N = 2 * (int64_t)c->n;
            n[r] = N;
            auto buf = std::make_unique<DBuf<double>>();
            buf->alloc((size_t)N * (5 + maxl + 1));
            double *b = buf->p;
            htab[(size_t)kSlotR * R + r] = b;
            htab[(size_t)kSlotZ * R + r] = b + N;
            htab[(size_t)kSlotWK * R + r] = b + 2 * N;
            htab[(size_t)kSlotDL * R + r] = b + 3 * N;
            htab[(size_t)kSlotT2 * R + r] = b + 4 * N;
            htab[(size_t)kSlotT1 * R + r] = c->vx.p;
            htab[(size_t)kSlotB * R + r] = dB[r];
            htab[(size_t)kSlotX * R + r] = dX[r];
            for (int k = 0; k <= maxl; ++k) htab[(size_t)(kSlotV0 + k) * R + r] = b + (size_t)(5 + k) * N;
            store.push_back(std::move(buf));
        }
        upload_tables();
        HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        join.resize(R, nullptr);
        for (auto &e : join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    ~MultiRuns() override
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (fork) (void)hipEventDestroy(fork);
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
    }
    // eager single-surface operations on the surfaces' own streams, ordered
    // after the common stream's work and before its next work
    void matvec(const std::vector<int32_t> &cols, const int32_t *, int sin, int sout) override
    {
        HIPCHK(hipStreamSynchronize(s));
        for (int r : cols) matvec_dev(cs[r], ptr(sin, r), ptr(sout, r));
        for (int r : cols) HIPCHK(hipStreamSynchronize(cs[r]->s));
    }
    void msolve(const std::vector<int32_t> &cols, const int32_t *, int sin, int sout) override
    {
        HIPCHK(hipStreamSynchronize(s));
        for (int r : cols) msolve_dev(cs[r], ptr(sin, r), ptr(sout, r));
        for (int r : cols) HIPCHK(hipStreamSynchronize(cs[r]->s));
    }
    // the step's work of one surface: vy = A vx, stage = M^-1 vy
    void step(int r) { matvec_dev(cs[r], cs[r]->vx.p, cs[r]->vy.p); msolve_dev(cs[r], cs[r]->vy.p, ptr(kSlotT2, r)); }
    hipGraphExec_t graph_for(const std::vector<int32_t> &cols)
    {
        auto it = graphs.find(cols);
        if (it != graphs.end()) return it->second;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        HIPCHK(hipEventRecord(fork, s));
        for (int r : cols) {
            HIPCHK(hipStreamWaitEvent(cs[r]->s, fork, 0));
            step(r);
            HIPCHK(hipEventRecord(join[r], cs[r]->s));
            HIPCHK(hipStreamWaitEvent(s, join[r], 0));
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        ++captures;
        graphs[cols] = ex;
        return ex;
    }
    void apply(const std::vector<int32_t> &cols, const int32_t *dcols, int sin, int sout) override
    {
        int64_t nmax = 0;
        for (int r : cols) nmax = std::max(nmax, n[r]);
        const dim3 grid(nblk(nmax, 256), (unsigned)cols.size());
        k_copyt<<<grid, 256, 0, s>>>(tab.p, R, sin, kSlotT1, dn.p, dcols);
        if (use_graph) {
            HIPCHK(hipGraphLaunch(graph_for(cols), s));
            ++graph_launches;
        } else {
            HIPCHK(hipEventRecord(fork, s));
            for (int r : cols) {
                HIPCHK(hipStreamWaitEvent(cs[r]->s, fork, 0));
                step(r);
                HIPCHK(hipEventRecord(join[r], cs[r]->s));
                HIPCHK(hipStreamWaitEvent(s, join[r], 0));
            }
        }
        k_copyt<<<grid, 256, 0, s>>>(tab.p, R, kSlotT2, sout, dn.p, dcols);
    }
};

// One GMRES run per surface of a packed context: every slot is one packed
// vector (2 * c->n doubles, padding zero), run r's vector is its surface's two
// segments in it, and one product or preconditioner solve serves every listed
// surface at once: the far, near and leaf-solve tasks of the other surfaces
// exit at once (the context's gate flags), leaving their entries stale, which
// nothing reads again (no interaction crosses surfaces).
struct PackedRuns : RunSet {
    tabi_ctx *c;
    DBuf<double> store;
    std::vector<double *> base;         // per slot: the packed vector
    DBuf<int32_t> dgate;
    std::vector<int32_t> hgate;         // flags last uploaded
    const int32_t *devgate = nullptr;   // device flags in use (graph form)
    bool device_gate_ok() const override { return true; }
    bool use_device_gate(const int32_t *g) override
    {
        devgate = g;
        c->gate = g;
        hgate.clear();
        return true;
    }
    void set_gate(const std::vector<int32_t> &cols)
    {
        if (devgate) { c->gate = devgate; return; }
        std::vector<int32_t> g(R, 0);
        for (int r : cols) g[r] = 1;
        if (g != hgate || !dgate.p) {
            hgate = g;
            dgate.upload(hgate.data(), (size_t)R, s);
        }
        c->gate = dgate.p;
    }
    ~PackedRuns() override { c->gate = nullptr; }
    PackedRuns(tabi_ctx *c_, int maxl_, double *dB, double *dX) : c(c_)
    {
        R = c->nsurf();
        maxl = maxl_;
        s = c->s;
        const int64_t L = 2 * (int64_t)c->n;
        for (int r = 0; r < R; ++r) maxl = (int)std::min<int64_t>(maxl, 2 * (int64_t)c->sf_n[r]);
        const int nown = nslots() - 2;  // every slot but B and X
        store.alloc((size_t)L * nown);
        HIPCHK(hipMemsetAsync(store.p, 0, sizeof(double) * (size_t)L * nown, s));
        base.assign(nslots(), nullptr);
        int k = 0;
        for (int sl = 0; sl < nslots(); ++sl) {
            if (sl == kSlotB) base[sl] = dB;
            else if (sl == kSlotX) base[sl] = dX;
            else base[sl] = store.p + (size_t)L * k++;
        }
        n.resize(R);
        gap.resize(R);
        htab.assign((size_t)nslots() * R, nullptr);
        for (int r = 0; r < R; ++r) {
            n[r] = 2 * (int64_t)c->sf_n[r];
            gap[r] = (int64_t)c->n - c->sf_n[r];
            for (int sl = 0; sl < nslots(); ++sl) htab[(size_t)sl * R + r] = base[sl] + c->sf_face0[r];
        }
        upload_tables();
    }
    void matvec(const std::vector<int32_t> &cols, const int32_t *, int sin, int sout) override
    {
        set_gate(cols);
        matvec_dev(c, base[sin], base[sout]);
    }
    void msolve(const std::vector<int32_t> &cols, const int32_t *, int sin, int sout) override
    {
        set_gate(cols);
        msolve_dev(c, base[sin], base[sout]);
    }
};

// GMRES with left preconditioning, ITOL=2, JSCAL=0, KMP=MAXL (DGMRES/DPIGMR,
// src/dgmres_dep.f:689-852,1652-1863): modified Gram-Schmidt with the
// conditional re-orthogonalisation of DORTH (1376-1411), Givens QR of the
// Hessenberg matrix (DHEQR, 1093-1125), residual vector for the stopping test
// and restarts from the Krylov basis (DRLCAL, 1948-1977), least squares update
// (DHELS, 925-947).  Vectors stay on the device; only scalars come back.
struct SingleRun : RunSet {
    tabi_ctx *c;
    DBuf<double> store;
    SingleRun(tabi_ctx *c_, int maxl_, const double *dB, double *dX) : c(c_)
    {
        R = 1;
        maxl = maxl_;
        s = c->s;
        const int64_t N = 2 * (int64_t)c->n;
        n.assign(1, N);
        store.alloc((size_t)N * (4 + maxl + 1));
        htab.assign(nslots(), nullptr);
        htab[kSlotR] = store.p;
        htab[kSlotZ] = store.p + N;
        htab[kSlotWK] = store.p + 2 * N;
        htab[kSlotDL] = store.p + 3 * N;
        htab[kSlotB] = const_cast<double *>(dB);   // read only
        htab[kSlotX] = dX;
        for (int k = 0; k <= maxl; ++k) htab[kSlotV0 + k] = store.p + (size_t)(4 + k) * N;
        upload_tables();
    }
    void matvec(const std::vector<int32_t> &, const int32_t *, int sin, int sout) override
    {
        matvec_dev(c, ptr(sin, 0), ptr(sout, 0));
    }
    void msolve(const std::vector<int32_t> &, const int32_t *, int sin, int sout) override
    {
        msolve_dev(c, ptr(sin, 0), ptr(sout, 0));
    }
    bool device_gate_ok() const override { return true; }
    bool use_device_gate(const int32_t *g) override
    {
        c->gate = g;
        return true;
    }
    ~SingleRun() override { c->gate = nullptr; }
};

// One GMRES run on the context's stream (gmres_runs with R = 1); with fallback,
// an initial guess whose preconditioned residual exceeds that of x0 = 0 is
// dropped and *x0_used reports whether the guess was kept.
void gmres_dev(tabi_ctx *c, const double *db, double *dxs, double tol, int maxl, int nrmax,
               int *iter_out, double *err_out, int *ierr_out, bool fallback = false,
               int *x0_used = nullptr, bool x0_zero = false)
{
    const int64_t N = 2 * (int64_t)c->n;
    if (maxl > N) maxl = (int)N;
    SingleRun rs(c, maxl, db, dxs);
    // flags bit 4: the cycles as captured graphs with the control on the device
    if ((c->prm.flags & 16) &&
        gmres_graph(rs, c->blk_red, tol, nrmax, x0_zero, iter_out, err_out, ierr_out, fallback, x0_used))
        return;
    gmres_runs(rs, c->blk_red, tol, nrmax, x0_zero, iter_out, err_out, ierr_out, fallback, x0_used);
}

// Right-hand sides of R charge sets (q[R][nq], host) at common positions
// into dB[R][2n].
void rhs_block_dev(tabi_ctx *c, int R, int nq, const double *qpos, const double *q, double eps_p,
                   double *dB)
{
    constexpr int RC = 8;
    DBuf<double> dqp, dq;
    dqp.upload(qpos, (size_t)3 * std::max(nq, 1), c->s);
    dq.upload(q, (size_t)R * std::max(nq, 1), c->s);
    if (c->prm.precision == 1)
        k_rhsb32<RC><<<dim3(nblk(c->n, 256), (R + RC - 1) / RC), 256, 0, c->s>>>(c->n, nq, R, dqp.p,
            dq.p, eps_p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, dB);
    else
        k_rhsb<RC><<<dim3(nblk(c->n, 256), (R + RC - 1) / RC), 256, 0, c->s>>>(c->n, nq, R, dqp.p, dq.p,
            eps_p, c->tx.p, c->ty.p, c->tz.p, c->tnx.p, c->tny.p, c->tnz.p, dB);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
}

// Energies of R solutions dX[R][2n] for the charge sets q[R][nq] (host).
void energy_block_dev(tabi_ctx *c, int R, int nq, const double *qpos, const double *q,
                      double eps_p, const double *dX, double *sol, double *cou)
{
    KTimer kt(c, 7);
    const double para = 332.0716;  // src/main.f90:64
    if (nq <= 0) {
        for (int r = 0; r < R; ++r) { sol[r] = 0.0; cou[r] = 0.0; }
        return;
    }
    constexpr int RC = 8;
    const int Rp = (R + RC - 1) / RC * RC;
    const int64_t N = 2 * (int64_t)c->n;
    DBuf<double> dqp, dq, dptl, part, wb, cpart;
    DBuf<int32_t> dall;
    std::vector<int32_t> all(R);
    for (int r = 0; r < R; ++r) all[r] = r;
    dall.upload(all.data(), R, c->s);
    dqp.upload(qpos, (size_t)3 * nq, c->s);
    dq.upload(q, (size_t)R * nq, c->s);
    const int ngq = (nq + kWave - 1) / kWave;
    const int nch = (c->n + kEnergyChunk - 1) / kEnergyChunk;
    part.alloc((size_t)ngq * nch * R * kWave);
    dptl.alloc((size_t)R * nq);
    if (c->prm.precision == 1) {
        // mixed precision: FP32 pair terms, as the single-solution energy
        DBuf<float> wbf;
        DBuf<float4> rec;
        wbf.alloc((size_t)c->n * 2 * Rp);
        rec.alloc((size_t)c->n * 2);
        k_src_weightsb<float><<<dim3(nblk(c->n, 256), Rp), 256, 0, c->s>>>(c->n, R, Rp, c->tarea.p, dX,
                                                                         N, dall.p, wbf.p);
        k_energy_rec32<<<nblk(c->n, 256), 256, 0, c->s>>>(c->n, c->src8.p, false, rec.p);
        k_energyb32<RC><<<dim3(ngq * nch, Rp / RC), 64, 0, c->s>>>(c->n, nq, nch, R, Rp, dqp.p,
            c->src8.p, rec.p, wbf.p, (float)c->prm.eps, (float)c->prm.kappa, part.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->s));
    } else {
        wb.alloc((size_t)c->n * 2 * Rp);
        k_src_weightsb<double><<<dim3(nblk(c->n, 256), Rp), 256, 0, c->s>>>(c->n, R, Rp, c->tarea.p, dX,
                                                                          N, dall.p, wb.p);
        k_energyb<RC><<<dim3(ngq * nch, Rp / RC), 64, 0, c->s>>>(c->n, nq, nch, R, Rp, dqp.p,
            c->src8.p, wb.p, c->prm.eps, c->prm.kappa, part.p);
    }
    k_energy_reduceb<<<dim3(nblk(nq, 256), R), 256, 0, c->s>>>(nq, nch, R, part.p, dptl.p);
    const int ncb = (nq + 255) / 256, ncp = ncb * ncb;
    cpart.alloc((size_t)R * ncp);
    for (int r = 0; r < R; ++r)
        k_coulomb<<<dim3(ncb, ncb), 256, 0, c->s>>>(nq, dqp.p, dq.p + (size_t)r * nq, cpart.p + (size_t)r * ncp);
    HIPCHK(hipGetLastError());
    std::vector<double> ptl((size_t)R * nq), cp((size_t)R * ncp);
    HIPCHK(hipMemcpyAsync(ptl.data(), dptl.p, sizeof(double) * ptl.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(cp.data(), cpart.p, sizeof(double) * cp.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (int r = 0; r < R; ++r) {
        double s = 0.0, cc = 0.0;
        for (int k = 0; k < nq; ++k) s += q[(size_t)r * nq + k] * ptl[(size_t)r * nq + k];
        for (int b = 0; b < ncp; ++b) cc += cp[(size_t)r * ncp + b];
        sol[r] = s * 0.5 * para * 4.0 * kPi;
        cou[r] = cc / eps_p * para;
    }
}

// Warm-start features of this surface for the charges (rhs db already formed);
// returns the charge scale s = 1/rms(b_1..b_n).
double ws_features_dev(tabi_ctx *c, int nq, const double *qpos, const double *q, double eps_p,
                       const double *db, float *dfeat, bool dev = false)
{
    const double bb = dev_dot(c, c->n, db, db);
    const double scale = bb > 0.0 ? 1.0 / std::sqrt(bb / c->n) : 1.0;
    DevIn dqp(qpos, (size_t)3 * nq, dev, c->s), dq(q, nq, dev, c->s);
    tabi::WsFeatIn in;
    in.n = c->n;
    in.tx = c->tx.p; in.ty = c->ty.p; in.tz = c->tz.p;
    in.tnx = c->tnx.p; in.tny = c->tny.p; in.tnz = c->tnz.p; in.tarea = c->tarea.p;
    in.leaf_of = c->leaf_of.p; in.leaf_beg = c->blk_ibeg.p; in.leaf_m2 = c->blk_m.p;
    in.nq = nq;
    in.qpos = dqp.p;
    in.q = dq.p;
    in.eps_p = eps_p;
    in.b = db;
    in.scale = scale;
    tabi::ws_features(in, dfeat, c->s);
    HIPCHK(hipStreamSynchronize(c->s));
    return scale;
}

// Scratch of the energy evaluation for nq charges on n faces: per (charge
// group, face chunk) partials, per-charge potentials, Coulomb partials.
// Mixed precision adds the FP32 source records (8 floats per face).
struct EnergyScratch {
    int ngq, nch, ncb, ncp;
    int64_t part, ptl, cpart, rec, total;   // offsets and size in doubles
    EnergyScratch(int n, int nq, bool mixed)
    {
        ngq = (nq + kWave - 1) / kWave;
        nch = (n + kEnergyChunk - 1) / kEnergyChunk;
        ncb = (nq + 255) / 256;
        ncp = ncb * ncb;
        part = 0;
        ptl = (int64_t)ngq * nch * kWave;
        cpart = ptl + nq;
        rec = (cpart + ncp + 1 + 1) / 2 * 2;   // 16-byte aligned
        total = mixed ? rec + 4 * (int64_t)n : cpart + ncp + 1;
    }
};

// Solvation and Coulomb energies (src/main.f90:144-157) from device inputs,
// with the scratch at w (EnergyScratch layout of nf faces); returns after
// completion.  Faces [f0, f0 + nf) of the context (default all); with
// weights false the source records already hold the solution's weights (a
// packed context forms them for every surface at once).
void energy_core(tabi_ctx *c, int nq, const double *dqp, const double *dq, double eps_p,
                 const double *dxs, double *w, double *sol, double *cou, int f0 = 0,
                 int nf = -1, bool weights = true)
{
    const double para = 332.0716;  // src/main.f90:64
    if (nf < 0) nf = c->n;
    const bool mixed = c->prm.precision == 1;
    const EnergyScratch L(nf, nq, mixed);
    double *part = w + L.part, *dptl = w + L.ptl, *cpart = w + L.cpart;
    if (weights) k_src_weights<<<nblk(c->n, 256), 256, 0, c->s>>>(c->n, c->tarea.p, dxs, c->src8.p);
    if (mixed) {
        float4 *rec = reinterpret_cast<float4 *>(w + L.rec);
        k_energy_rec32<<<nblk(nf, 256), 256, 0, c->s>>>(nf, c->src8.p + (size_t)8 * f0, true, rec);
        k_energy32<<<L.ngq * L.nch, 64, 0, c->s>>>(nf, nq, L.nch, dqp, c->src8.p + (size_t)8 * f0,
                                                   rec, (float)c->prm.eps, (float)c->prm.kappa, part);
    } else {
        k_energy<<<L.ngq * L.nch, 64, 0, c->s>>>(nf, nq, L.nch, dqp, c->src8.p + (size_t)8 * f0,
                                                 c->prm.eps, c->prm.kappa, part);
    }
    k_energy_reduce<<<nblk(nq, 256), 256, 0, c->s>>>(nq, L.nch, part, dptl);
    k_coulomb<<<dim3(L.ncb, L.ncb), 256, 0, c->s>>>(nq, dqp, dq, cpart);
    k_sum_part<<<1, 256, 0, c->s>>>(L.ncp, cpart, cpart + L.ncp);
    HIPCHK(hipGetLastError());
    // sum_k q_k * potential_k (src/main.f90:148,156)
    const double sdot = dev_dot(c, nq, dq, dptl);
    double cc = 0.0;
    HIPCHK(hipMemcpyAsync(&cc, cpart + L.ncp, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    *sol = sdot * 0.5 * para * 4.0 * kPi;
    *cou = cc / eps_p * para;
}

void energy_dev(tabi_ctx *c, int nq, const double *qpos, const double *q, double eps_p,
                const double *dxs, double *sol, double *cou, bool dev = false)
{
    KTimer kt(c, 7);
    if (nq <= 0) { *sol = 0.0; *cou = 0.0; return; }
    DevIn dqp(qpos, (size_t)3 * nq, dev, c->s), dq(q, nq, dev, c->s);
    DBuf<double> w;
    w.alloc((size_t)EnergyScratch(c->n, nq, c->prm.precision == 1).total);
    energy_core(c, nq, dqp.p, dq.p, eps_p, dxs, w.p, sol, cou);
}

// ---------------------------------------------------------------------------
// Allocation-free energy (tabi_ctx_energy_noalloc): the scratch lives in a
// file-scope device array of fixed size, one copy per GPU, whose address is
// looked up once per device.  Host code uses only that address (the array's
// name on the host is not a device pointer).  Every call on a GPU shares the
// array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int64_t kScratchDoubles = (int64_t)4 << 20;     // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;

double *scratch_ptr()
{
    static std::mutex mu;
    static double *ptrs[kMaxDevices] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("device index out of range");
    std::lock_guard<std::mutex> g(mu);
    if (!ptrs[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptrs[dev] = static_cast<double *>(p);
    }
    return ptrs[dev];
}

// ---------------------------------------------------------------------------
// Vertex-averaged surface potential (output_potential, src/main.f90:226-326).
// Every vertex takes the first 14 faces that use it, in the original face
// numbering (the reference's ind_vert slots, nface_vert-1 = 14; later faces are
// dropped), and averages phi and dphi/dn over them weighted by 1/|centroid -
// vertex|; the results are scaled by para*4*pi as the file holds them.  A vertex
// that no face uses gets 0/0 = NaN, as in the reference.  The incidences
// (vertex, face) are listed in face order and stably sorted by vertex, which
// keeps each vertex's faces ascending.
// ---------------------------------------------------------------------------
constexpr int kVertFaces = 14;

__global__ void k_vp_incid(int n, int nvert, const int32_t *__restrict__ faces1,
                           int32_t *__restrict__ key, int32_t *__restrict__ val,
                           int32_t *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * (int64_t)n) return;
    int v = faces1[i] - 1;
    if (v < 0 || v >= nvert) {
        atomicOr(bad, 1);
        v = nvert;          // sorted past every vertex, never read
    }
    key[i] = v;
    val[i] = (int32_t)(i / 3);
}

// beg[v] = first sorted incidence of vertex v (lower bound), v = 0..nvert
__global__ void k_vp_seg(int m, int nvert, const int32_t *__restrict__ key,
                         int32_t *__restrict__ beg)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > nvert) return;
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] < v) lo = mid + 1; else hi = mid;
    }
    beg[v] = lo;
}

__global__ void k_vertex_pot(int n, int nvert, const int32_t *__restrict__ beg,
                             const int32_t *__restrict__ face, const int32_t *__restrict__ inv,
                             const double *__restrict__ vpos, const double *__restrict__ tx,
                             const double *__restrict__ ty, const double *__restrict__ tz,
                             const double *__restrict__ x, double scale, double *__restrict__ out)
{
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvert) return;
    const int b = beg[v], e = min(beg[v + 1], b + kVertFaces);
    const double px = vpos[3 * (int64_t)v], py = vpos[3 * (int64_t)v + 1], pz = vpos[3 * (int64_t)v + 2];
    double s1 = 0.0, s2 = 0.0, tot = 0.0;
    for (int k = b; k < e; ++k) {
        const int t = inv[face[k]];
        const double ax = tx[t] - px, ay = ty[t] - py, az = tz[t] - pz;
        const double w = 1.0 / sqrt(ax * ax + ay * ay + az * az);
        s1 = s1 + w * x[t];
        s2 = s2 + w * x[n + t];
        tot = tot + w;
    }
    out[2 * (int64_t)v] = s1 / tot * scale;
    out[2 * (int64_t)v + 1] = s2 / tot * scale;
}

// faces1[3n]: the surface's faces in the original numbering (1-based vertex
// indices, the nvert array of the reference after culling); x: tree-order
// solution; vpot[2*nvert]: (phi, dphi/dn) per vertex.  Host pointers.
void vertex_potential(tabi_ctx *c, int nvert, const double *vpos, const int32_t *faces1,
                      const double *x, double *vpot)
{
    if (nvert < 1) throw Error("need at least one vertex");
    const int n = c->n;
    const int m = 3 * n;
    DBuf<int32_t> dfaces, key, val, key2, val2, beg, inv, bad;
    DBuf<double> dv, dx, dout;
    dfaces.upload(faces1, (size_t)m, c->s);
    dv.upload(vpos, (size_t)3 * nvert, c->s);
    dx.upload(x, (size_t)2 * n, c->s);
    {
        std::vector<int32_t> hinv(n);
        for (int t = 0; t < n; ++t) hinv[c->tree.order[t]] = t;
        inv.upload(hinv.data(), n, c->s);
    }
    key.alloc(m); val.alloc(m); key2.alloc(m); val2.alloc(m);
    beg.alloc((size_t)nvert + 1); bad.alloc(1); dout.alloc((size_t)2 * nvert);
    HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), c->s));
    k_vp_incid<<<nblk(m, 256), 256, 0, c->s>>>(n, nvert, dfaces.p, key.p, val.p, bad.p);
    HIPCHK(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 31 && (1 << end_bit) <= nvert) ++end_bit;
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key2.p, val.p, val2.p, m, 0,
                                              end_bit, c->s));
    DBuf<char> tmp;
    tmp.alloc(std::max<size_t>(tmp_bytes, 1));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key2.p, val.p, val2.p, m, 0,
                                              end_bit, c->s));
    k_vp_seg<<<nblk((int64_t)nvert + 1, 256), 256, 0, c->s>>>(m, nvert, key2.p, beg.p);
    const double scale = 332.0716 * 4.0 * kPi;   // para*4*pi, src/main.f90:64,245
    k_vertex_pot<<<nblk(nvert, 256), 256, 0, c->s>>>(n, nvert, beg.p, val2.p, inv.p, dv.p, c->tx.p,
                                                     c->ty.p, c->tz.p, dx.p, scale, dout.p);
    HIPCHK(hipGetLastError());
    int32_t hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(vpot, dout.p, sizeof(double) * 2 * nvert, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (hbad) throw Error("face vertex index outside 1.." + std::to_string(nvert));
}

// One value in a Fortran Fw.d field: fixed notation right-aligned in w
// columns, asterisks when it does not fit, NaN / Infinity spelled out.
void fortran_f(std::string &out, double v, int w, int d)
{
    char buf[64];
    if (std::isnan(v)) std::snprintf(buf, sizeof buf, "NaN");
    else if (std::isinf(v)) std::snprintf(buf, sizeof buf, v > 0 ? "Infinity" : "-Infinity");
    else std::snprintf(buf, sizeof buf, "%.*f", d, v);
    const int len = (int)std::strlen(buf);
    if (len > w) { out.append((size_t)w, '*'); return; }
    out.append((size_t)(w - len), ' ');
    out.append(buf);
}

std::unique_ptr<tabi_ctx> g_default;

template <typename F>
int guarded(F &&f)
{
    try {
        g_err.clear();
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

tabi_ctx *need_default()
{
    if (!g_default) throw Error("tabi_init has not been called");
    return g_default.get();
}

void check_n2(tabi_ctx *c, int32_t n2)
{
    if (n2 != 2 * c->n) throw Error("vector length " + std::to_string(n2) + " does not match 2*nface = " +
                                    std::to_string(2 * c->n));
}

void ctx_matvec_host(tabi_ctx *c, const double *x, double *y)
{
    const int64_t N = 2 * (int64_t)c->n;
    double sabs = 0.0;
    for (int64_t i = 0; i < N; ++i) sabs += std::fabs(x[i]);
    if (sabs < 1.0e-10) {  // src/main.f90:478-479
        std::memcpy(y, x, sizeof(double) * N);
        return;
    }
    HIPCHK(hipMemcpyAsync(c->vx.p, x, sizeof(double) * N, hipMemcpyHostToDevice, c->s));
    matvec_dev(c, c->vx.p, c->vy.p);
    HIPCHK(hipMemcpyAsync(y, c->vy.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
}

void ctx_msolve_host(tabi_ctx *c, const double *r, double *z)
{
    const int64_t N = 2 * (int64_t)c->n;
    DBuf<double> dr;
    dr.upload(r, N, c->s);
    msolve_dev(c, dr.p, c->vy.p);
    HIPCHK(hipMemcpyAsync(z, c->vy.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
}

void ctx_gmres_host(tabi_ctx *c, const double *b, double *x, double tol, int maxl, int nrmax,
                    int *iter, double *err, int *ierr)
{
    const int64_t N = 2 * (int64_t)c->n;
    DBuf<double> db, dxs;
    db.upload(b, N, c->s);
    dxs.upload(x, N, c->s);
    gmres_dev(c, db.p, dxs.p, tol, maxl, nrmax, iter, err, ierr);
    HIPCHK(hipMemcpyAsync(x, dxs.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int tabi_init(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
              const double *area, int32_t *orderarr_out, int32_t *nleaf_out)
{
    return guarded([&] {
        auto c = std::make_unique<tabi_ctx>();
        ctx_setup(c.get(), prm, n, xyz, nrm, area, orderarr_out, nullptr);
        if (nleaf_out) *nleaf_out = c->nleaf;
        if (g_default) {
            hipStream_t s0 = g_default->s;
            g_default->s = nullptr;
            destroy_stream(s0);
        }
        g_default = std::move(c);
    });
}

int tabi_matvec(int32_t n2, const double *x, double *y)
{
    return guarded([&] {
        tabi_ctx *c = need_default();
        PoolScope ps_(c->s);
        check_n2(c, n2);
        ctx_matvec_host(c, x, y);
    });
}

int tabi_msolve(int32_t n2, const double *r, double *z)
{
    return guarded([&] {
        tabi_ctx *c = need_default();
        PoolScope ps_(c->s);
        check_n2(c, n2);
        ctx_msolve_host(c, r, z);
    });
}

int tabi_rhs(int32_t nq, const double *qpos, const double *q, double eps_p, double *b)
{
    return tabi_ctx_rhs(g_default.get(), nq, qpos, q, eps_p, b);
}

int tabi_energy(int32_t nq, const double *qpos, const double *q, double eps_p, const double *x,
                double *soleng, double *couleng)
{
    return tabi_ctx_energy(g_default.get(), nq, qpos, q, eps_p, x, soleng, couleng);
}

int tabi_gmres(const double *b, double *x, double tol, int32_t maxl, int32_t nrmax,
               int32_t *iter, double *err, int32_t *ierr)
{
    return tabi_ctx_gmres(g_default.get(), b, x, tol, maxl, nrmax, iter, err, ierr);
}

void tabi_finalize(void)
{
    if (g_default) {
        hipStream_t s = g_default->s;
        if (s) (void)hipStreamSynchronize(s);
        g_default.reset();
        destroy_stream(s);
    }
}

const char *tabi_last_error(void) { return g_err.c_str(); }

int tabi_stats(int64_t *out10)
{
    return guarded([&] { ctx_stats(need_default(), out10); });
}

int tabi_ctx_create(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
                    const double *area, int32_t *orderarr_out, tabi_ctx **out)
{
    return guarded([&] {
        auto c = std::make_unique<tabi_ctx>();
        ctx_setup(c.get(), prm, n, xyz, nrm, area, orderarr_out, nullptr);
        *out = c.release();
    });
}

int tabi_ctx_create_packed(const tabi_params *prm, int32_t nsurf, const int32_t *nface,
                           const double *const *xyz, const double *const *nrm,
                           const double *const *area, int32_t *face0_out, tabi_ctx **out)
{
    return guarded([&] {
        auto c = std::make_unique<tabi_ctx>();
        ctx_setup_packed(c.get(), prm, nsurf, nface, xyz, nrm, area);
        if (face0_out) std::copy(c->sf_face0.begin(), c->sf_face0.end(), face0_out);
        *out = c.release();
    });
}

void tabi_ctx_destroy(tabi_ctx *c)
{
    if (!c) return;
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    hipStream_t s = c->s;
    if (s) (void)hipStreamSynchronize(s);  // pending work may still use the buffers
    delete c;
    destroy_stream(s);
}

int tabi_ctx_stats(tabi_ctx *c, int64_t *out10)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        ctx_stats(c, out10);
    });
}

int tabi_ctx_matvec_dev(tabi_ctx *c, const double *d_x, double *d_y)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        matvec_dev(c, d_x, d_y);
    });
}

int tabi_ctx_msolve_dev(tabi_ctx *c, const double *d_r, double *d_z)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        msolve_dev(c, d_r, d_z);
    });
}

int tabi_ctx_matvec_block(tabi_ctx *c, int32_t nrhs, const double *x, double *y)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (nrhs < 1) throw Error("need at least one right-hand side");
        const int64_t N = 2 * (int64_t)c->n;
        std::vector<int32_t> live;
        for (int r = 0; r < nrhs; ++r) {
            double sabs = 0.0;
            for (int64_t i = 0; i < N; ++i) sabs += std::fabs(x[(size_t)r * N + i]);
            if (sabs < 1.0e-10)  // src/main.f90:478-479, per vector
                std::memcpy(y + (size_t)r * N, x + (size_t)r * N, sizeof(double) * N);
            else
                live.push_back(r);
        }
        if (live.empty()) return;
        DBuf<double> dX, dY;
        DBuf<int32_t> dact;
        dX.upload(x, (size_t)nrhs * N, c->s);
        dY.alloc((size_t)nrhs * N);
        dact.upload(live.data(), live.size(), c->s);
        matvec_block_dev(c, (int)live.size(), dX.p, N, dY.p, N, dact.p);
        std::vector<double> hy((size_t)nrhs * N);
        HIPCHK(hipMemcpyAsync(hy.data(), dY.p, sizeof(double) * hy.size(), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        for (int r : live) std::memcpy(y + (size_t)r * N, hy.data() + (size_t)r * N, sizeof(double) * N);
    });
}

int tabi_ctx_msolve_block(tabi_ctx *c, int32_t nrhs, const double *r, double *z)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (nrhs < 1) throw Error("need at least one right-hand side");
        const int64_t N = 2 * (int64_t)c->n;
        DBuf<double> dR, dZ;
        DBuf<int32_t> dact;
        std::vector<int32_t> all(nrhs);
        for (int k = 0; k < nrhs; ++k) all[k] = k;
        dR.upload(r, (size_t)nrhs * N, c->s);
        dZ.alloc((size_t)nrhs * N);
        dact.upload(all.data(), nrhs, c->s);
        msolve_block_dev(c, nrhs, dR.p, N, dZ.p, N, dact.p);
        HIPCHK(hipMemcpyAsync(z, dZ.p, sizeof(double) * nrhs * N, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
    });
}

int tabi_ctx_gmres_block(tabi_ctx *c, int32_t nrhs, const double *b, double *x, double tol,
                         int32_t maxl, int32_t nrmax, int32_t *iter, double *err, int32_t *ierr)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (nrhs < 1) throw Error("need at least one right-hand side");
        if (maxl < 1) throw Error("maxl must be >= 1");
        const int64_t N = 2 * (int64_t)c->n;
        DBuf<double> dB, dX;
        dB.upload(b, (size_t)nrhs * N, c->s);
        dX.upload(x, (size_t)nrhs * N, c->s);
        std::vector<int> it(nrhs), ie(nrhs);
        gmres_block_dev(c, nrhs, dB.p, dX.p, tol, maxl, nrmax, false, it.data(), err, ie.data());
        HIPCHK(hipMemcpyAsync(x, dX.p, sizeof(double) * nrhs * N, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        for (int k = 0; k < nrhs; ++k) { iter[k] = it[k]; ierr[k] = ie[k]; }
    });
}

int tabi_ctx_solve_block(tabi_ctx *c, int32_t nrhs, int32_t nq, const double *qpos,
                         const double *q, double eps_p, double tol, double *x_out, int32_t *iter,
                         double *err, int32_t *ierr, double *soleng, double *couleng)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (nrhs < 1) throw Error("need at least one charge set");
        if (nq < 1) throw Error("need at least one charge");
        const int64_t N = 2 * (int64_t)c->n;
        DBuf<double> dB, dX;
        dB.alloc((size_t)nrhs * N);
        dX.alloc((size_t)nrhs * N);
        rhs_block_dev(c, nrhs, nq, qpos, q, eps_p, dB.p);
        HIPCHK(hipMemsetAsync(dX.p, 0, sizeof(double) * nrhs * N, c->s));
        std::vector<int> it(nrhs), ie(nrhs);
        gmres_block_dev(c, nrhs, dB.p, dX.p, tol, 10, 10, true, it.data(), err, ie.data());
        energy_block_dev(c, nrhs, nq, qpos, q, eps_p, dX.p, soleng, couleng);
        if (x_out) {
            HIPCHK(hipMemcpyAsync(x_out, dX.p, sizeof(double) * nrhs * N, hipMemcpyDeviceToHost, c->s));
            HIPCHK(hipStreamSynchronize(c->s));
        }
        for (int k = 0; k < nrhs; ++k) { iter[k] = it[k]; ierr[k] = ie[k]; }
    });
}

// Many surfaces packed into one context (ctx_setup_packed): one set-up (the
// trees from host threads, then one list build and one batched leaf LU over
// every surface), right-hand sides per surface, the GMRES runs of all surfaces
// in lockstep over the packed operator, energies per surface.
void solve_many_packed(const tabi_params *prm, int32_t nsurf, const int32_t *nface,
                       const double *const *xyz, const double *const *nrm,
                       const double *const *area, const int32_t *nq, const double *const *qpos,
                       const double *const *q, double eps_p, double tol, int32_t options,
                       int32_t *iter, double *err, int32_t *ierr, double *soleng, double *couleng)
{
    (void)options;
    auto t0 = std::chrono::steady_clock::now();
    tabi_ctx c;
    struct StreamGuard {
        tabi_ctx *c;
        ~StreamGuard()
        {
            if (c->s) { hipStream_t st = c->s; c->s = nullptr; destroy_stream(st); }
        }
    } guard{&c};
    ctx_setup_packed(&c, prm, nsurf, nface, xyz, nrm, area, g_many + 6);
    auto tr = std::chrono::steady_clock::now();
    PoolScope ps_(c.s);
    const int64_t L = 2 * (int64_t)c.n;
    DBuf<double> bx;
    bx.alloc((size_t)2 * L);
    double *dB = bx.p, *dX = bx.p + L;
    HIPCHK(hipMemsetAsync(bx.p, 0, sizeof(double) * 2 * L, c.s));
    for (int k = 0; k < nsurf; ++k)
        rhs_dev(&c, nq[k], qpos[k], q[k], eps_p, dB, false, c.sf_face0[k], nface[k], false);
    HIPCHK(hipStreamSynchronize(c.s));
    auto t1 = std::chrono::steady_clock::now();
    std::vector<int> it(nsurf), ie(nsurf);
    {
        PackedRuns rs(&c, 10, dB, dX);
        int64_t launches = 0;
        if (!(options & 1) &&
            gmres_graph(rs, c.blk_red, tol, 10, true, it.data(), err, ie.data(), false, nullptr, &launches)) {
            g_many[0] = 1.0;                 // one cycle graph, captured once
            g_many[1] = (double)launches;    // launched once per restart cycle
        } else {
            gmres_runs(rs, c.blk_red, tol, 10, true, it.data(), err, ie.data());
            g_many[0] = 0.0;
            g_many[1] = 0.0;
        }
    }
    HIPCHK(hipStreamSynchronize(c.s));
    auto t2 = std::chrono::steady_clock::now();
    k_src_weights<<<nblk(c.n, 256), 256, 0, c.s>>>(c.n, c.tarea.p, dX, c.src8.p);
    HIPCHK(hipGetLastError());
    for (int k = 0; k < nsurf; ++k) {
        DevIn dqp(qpos[k], (size_t)3 * nq[k], false, c.s), dq(q[k], nq[k], false, c.s);
        DBuf<double> w;
        w.alloc((size_t)EnergyScratch(nface[k], nq[k], c.prm.precision == 1).total);
        energy_core(&c, nq[k], dqp.p, dq.p, eps_p, dX, w.p, soleng + k, couleng + k, c.sf_face0[k],
                    nface[k], false);
        iter[k] = it[k];
        ierr[k] = ie[k];
    }
    auto t3 = std::chrono::steady_clock::now();
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    g_many[2] = ms(t0, t1);
    g_many[3] = ms(t1, t2);
    g_many[4] = ms(t2, t3);
    g_many[5] = ms(t0, t3);
    g_many[9] = ms(tr, t1);
}

int tabi_solve_many(const tabi_params *prm, int32_t nsurf, const int32_t *nface,
                    const double *const *xyz, const double *const *nrm,
                    const double *const *area, const int32_t *nq, const double *const *qpos,
                    const double *const *q, double eps_p, double tol, int32_t options,
                    int32_t *iter, double *err, int32_t *ierr, double *soleng, double *couleng)
{
    return guarded([&] {
        if (nsurf < 1) throw Error("need at least one surface");
        for (int k = 0; k < nsurf; ++k)
            if (nq[k] < 1) throw Error("surface " + std::to_string(k) + " has no charges");
        if (!(options & 2)) {
            solve_many_packed(prm, nsurf, nface, xyz, nrm, area, nq, qpos, q, eps_p, tol, options,
                              iter, err, ierr, soleng, couleng);
            return;
        }
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::unique_ptr<tabi_ctx>> ctx(nsurf);
        std::vector<std::unique_ptr<DBuf<double>>> bx(nsurf);
        std::vector<double *> dB(nsurf), dX(nsurf);
        for (int k = 0; k < nsurf; ++k) {
            ctx[k] = std::make_unique<tabi_ctx>();
            bx[k] = std::make_unique<DBuf<double>>();
        }
        auto cleanup = [&]() {
            for (auto &c : ctx)
                if (c && c->s) {
                    hipStream_t st = c->s;
                    c->s = nullptr;
                    destroy_stream(st);
                }
        };
        // set-up and right-hand sides, several surfaces at a time from host threads
        {
            const int nth = std::max(1, std::min<int>(nsurf, 8));
            std::vector<std::thread> th;
            std::vector<std::string> errs(nth);
            for (int t = 0; t < nth; ++t)
                th.emplace_back([&, t] {
                    try {
                        for (int k = t; k < nsurf; k += nth) {
                            tabi_ctx *c = ctx[k].get();
                            ctx_setup(c, prm, nface[k], xyz[k], nrm[k], area[k], nullptr, nullptr);
                            PoolScope ps_(c->s);
                            const int64_t N = 2 * (int64_t)nface[k];
                            bx[k]->alloc((size_t)2 * N);
                            rhs_dev(c, nq[k], qpos[k], q[k], eps_p, bx[k]->p);
                            HIPCHK(hipMemsetAsync(bx[k]->p + N, 0, sizeof(double) * N, c->s));
                            HIPCHK(hipStreamSynchronize(c->s));
                        }
                    } catch (const std::exception &e) {
                        errs[t] = e.what();
                    }
                });
            for (auto &x : th) x.join();
            for (auto &e : errs)
                if (!e.empty()) { cleanup(); throw Error(e); }
        }
        for (int k = 0; k < nsurf; ++k) {
            dB[k] = bx[k]->p;
            dX[k] = bx[k]->p + 2 * (int64_t)nface[k];
        }
        auto t1 = std::chrono::steady_clock::now();
        hipStream_t main = nullptr;
        HIPCHK(hipStreamCreate(&main));
        std::vector<int> it(nsurf), ie(nsurf);
        try {
            std::vector<tabi_ctx *> cp(nsurf);
            for (int k = 0; k < nsurf; ++k) cp[k] = ctx[k].get();
            DBuf<double> red;
            {
                MultiRuns rs(cp, 10, main, (options & 1) == 0, dB, dX);
                gmres_runs(rs, red, tol, 10, true, it.data(), err, ie.data());
                g_many[0] = (double)rs.captures;
                g_many[1] = (double)rs.graph_launches;
            }
        } catch (...) {
            destroy_stream(main);
            cleanup();
            throw;
        }
        HIPCHK(hipStreamSynchronize(main));
        destroy_stream(main);
        auto t2 = std::chrono::steady_clock::now();
        for (int k = 0; k < nsurf; ++k) {
            PoolScope ps_(ctx[k]->s);
            energy_dev(ctx[k].get(), nq[k], qpos[k], q[k], eps_p, dX[k], soleng + k, couleng + k);
            iter[k] = it[k];
            ierr[k] = ie[k];
        }
        auto t3 = std::chrono::steady_clock::now();
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        g_many[2] = ms(t0, t1);
        g_many[3] = ms(t1, t2);
        g_many[4] = ms(t2, t3);
        g_many[5] = ms(t0, t3);
        cleanup();
    });
}

int tabi_last_solve_many_stats(double *out6)
{
    for (int i = 0; i < 6; ++i) out6[i] = g_many[i];
    return 0;
}

int tabi_last_solve_many_setup(double *out4)
{
    for (int i = 0; i < 4; ++i) out4[i] = g_many[6 + i];
    return 0;
}

int tabi_release_memory(void)
{
    return guarded([&] { DevicePool::get().release_all(); });
}

int tabi_device_alloc(int64_t bytes, void **out)
{
    return guarded([&] {
        if (bytes <= 0) throw Error("allocation size must be positive");
        HIPCHK(hipMalloc(out, (size_t)bytes));
    });
}

int tabi_device_free(void *p)
{
    return guarded([&] { HIPCHK(hipFree(p)); });
}

int tabi_copy_to_device(void *dst, const void *src, int64_t bytes)
{
    return guarded([&] { HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice)); });
}

int tabi_copy_to_host(void *dst, const void *src, int64_t bytes)
{
    return guarded([&] { HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost)); });
}

int tabi_ctx_sync(tabi_ctx *c)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        HIPCHK(hipStreamSynchronize(c->s));
    });
}

int tabi_ctx_matvec(tabi_ctx *c, const double *x, double *y)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        ctx_matvec_host(c, x, y);
    });
}

int tabi_ctx_msolve(tabi_ctx *c, const double *r, double *z)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        ctx_msolve_host(c, r, z);
    });
}

int tabi_ctx_rhs(tabi_ctx *c, int32_t nq, const double *qpos, const double *q, double eps_p,
                 double *b)
{
    return guarded([&] {
        if (!c) throw Error("tabi_init has not been called");
        PoolScope ps_(c->s);
        DBuf<double> db;
        db.alloc((size_t)2 * c->n);
        rhs_dev(c, nq, qpos, q, eps_p, db.p);
        HIPCHK(hipMemcpyAsync(b, db.p, sizeof(double) * 2 * c->n, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
    });
}

int tabi_ctx_energy(tabi_ctx *c, int32_t nq, const double *qpos, const double *q, double eps_p,
                    const double *x, double *soleng, double *couleng)
{
    return guarded([&] {
        if (!c) throw Error("tabi_init has not been called");
        PoolScope ps_(c->s);
        DBuf<double> dxs;
        dxs.upload(x, (size_t)2 * c->n, c->s);
        energy_dev(c, nq, qpos, q, eps_p, dxs.p, soleng, couleng);
    });
}

int tabi_ctx_vertex_potential(tabi_ctx *c, int32_t nvert, const double *vpos,
                              const int32_t *faces, const double *x, double *vpot)
{
    return guarded([&] {
        if (!c) throw Error("tabi_init has not been called");
        PoolScope ps_(c->s);
        if (!vpos || !faces || !x || !vpot) throw Error("null array");
        vertex_potential(c, nvert, vpos, faces, x, vpot);
    });
}

int tabi_vertex_potential(int32_t nvert, const double *vpos, const int32_t *faces, const double *x,
                          double *vpot)
{
    return tabi_ctx_vertex_potential(g_default.get(), nvert, vpos, faces, x, vpot);
}

int tabi_write_surface_potential(const char *path, int32_t nvert, const double *vpos,
                                 const double *vnrm, const double *vpot, int32_t nface,
                                 const int32_t *faces)
{
    return guarded([&] {
        if (!path || nvert < 0 || nface < 0) throw Error("bad arguments");
        FILE *f = std::fopen(path, "w");
        if (!f) throw Error(std::string("cannot open ") + path);
        std::string line;
        line.reserve(160);
        bool ok = std::fprintf(f, " %d %d\n", nvert, nface) > 0;
        for (int64_t i = 0; ok && i < nvert; ++i) {
            char idx[32];
            std::snprintf(idx, sizeof idx, "%10lld", (long long)(i + 1));
            line.assign(idx);
            for (int k = 0; k < 3; ++k) fortran_f(line, vpos[3 * i + k], 12, 6);
            for (int k = 0; k < 3; ++k) fortran_f(line, vnrm[3 * i + k], 12, 6);
            for (int k = 0; k < 2; ++k) fortran_f(line, vpot[2 * i + k], 20, 10);
            line.push_back('\n');
            ok = std::fputs(line.c_str(), f) >= 0;
        }
        for (int64_t j = 0; ok && j < nface; ++j)
            ok = std::fprintf(f, " %d %d %d\n", faces[3 * j], faces[3 * j + 1], faces[3 * j + 2]) > 0;
        if (std::fclose(f) != 0 || !ok) throw Error(std::string("write failed: ") + path);
    });
}

int tabi_ctx_energy_noalloc(tabi_ctx *c, int32_t nq, const double *d_qpos, const double *d_q,
                            double eps_p, const double *d_x, double *soleng, double *couleng)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        if (!soleng || !couleng) throw Error("null output");
        if (nq <= 0) { *soleng = 0.0; *couleng = 0.0; return; }
        if (!d_qpos || !d_q || !d_x) throw Error("null device array");
        const EnergyScratch L(c->n, nq, c->prm.precision == 1);
        if (L.total > kScratchDoubles)
            throw Error("tabi_ctx_energy_noalloc: " + std::to_string(nq) + " charges on " +
                        std::to_string(c->n) + " faces need " + std::to_string(L.total * 8) +
                        " bytes of scratch, more than the static " +
                        std::to_string(kScratchDoubles * 8) + " (use tabi_ctx_energy)");
        double *w = scratch_ptr();
        energy_core(c, nq, d_qpos, d_q, eps_p, d_x, w, soleng, couleng);
    });
}

int tabi_ctx_gmres(tabi_ctx *c, const double *b, double *x, double tol, int32_t maxl,
                   int32_t nrmax, int32_t *iter, double *err, int32_t *ierr)
{
    return guarded([&] {
        if (!c) throw Error("tabi_init has not been called");
        PoolScope ps_(c->s);
        int it = 0, ie = 0;
        ctx_gmres_host(c, b, x, tol, maxl, nrmax, &it, err, &ie);
        *iter = it;
        *ierr = ie;
    });
}

int tabi_solve(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
               const double *area, int32_t nq, const double *qpos, const double *q,
               double eps_p, double tol, double *x_out, int32_t *orderarr_out, int32_t *iter,
               double *err, int32_t *ierr, double *soleng, double *couleng)
{
    return tabi_solve_ws(prm, n, xyz, nrm, area, nq, qpos, q, eps_p, tol, nullptr, x_out,
                         orderarr_out, iter, err, ierr, soleng, couleng, nullptr);
}

}  // extern "C"

namespace {

// One full solve; with dev the geometry, the charges and x_out are device
// pointers (inputs already resident), otherwise host pointers.
void solve_core(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
                const double *area, int32_t nq, const double *qpos, const double *q, double eps_p,
                double tol, tabi_ws *model, double *x_out, int32_t *orderarr_out, int32_t *iter,
                double *err, int32_t *ierr, double *soleng, double *couleng, int32_t *ws_used,
                bool dev)
{
    if (model && !model->m) throw Error("null warm-start model");
    if (nq < 1) throw Error("need at least one charge");
    auto t0 = std::chrono::steady_clock::now();
    tabi_ctx c;
    hipStream_t s2 = nullptr;
    try {
        double ph[3] = {0, 0, 0};
        const bool split = !(prm && (prm->flags & 4));
        c.s = StreamPool::acquire();
        s2 = StreamPool::acquire();
        ctx_setup(&c, prm, n, xyz, nrm, area, orderarr_out, ph, dev, !split);
        PoolScope ps_(c.s);
        const int64_t N = 2 * (int64_t)n;
        DBuf<double> db, dxs_own;
        db.alloc(N);
        double *dxs = (dev && x_out) ? x_out : (dxs_own.alloc(N), dxs_own.p);
        HIPCHK(hipMemsetAsync(dxs, 0, sizeof(double) * N, c.s));
        // The right-hand side and the learned guess need the tree and the
        // leaf tables but not the factors: they run on a second stream beside
        // the factorisation (whose panel steps leave most of the GPU idle).
        EventSet ev(2);
        HIPCHK(hipEventRecord(ev.e[0], c.s));
        HIPCHK(hipStreamWaitEvent(s2, ev.e[0], 0));
        auto t1 = std::chrono::steady_clock::now();
        if (split) precond_factor(&c);
        const hipStream_t s_main = c.s;
        double t_rhs = 0.0, t_ws = 0.0;
        {
            // the helpers below queue on the context's stream: point it at s2
            c.s = s2;
            PoolScope ps2(s2);
            try {
                auto a = std::chrono::steady_clock::now();
                rhs_dev(&c, nq, qpos, q, eps_p, db.p, dev);
                auto b = std::chrono::steady_clock::now();
                if (model) {
                    DBuf<float> feat;
                    feat.alloc((size_t)n * tabi::ws::kF);
                    const double scale = ws_features_dev(&c, nq, qpos, q, eps_p, db.p, feat.p, dev);
                    tabi::ws_predict_dev(model->m, n, feat.p, scale, dxs, s2);
                    HIPCHK(hipStreamSynchronize(s2));
                }
                auto d = std::chrono::steady_clock::now();
                t_rhs = std::chrono::duration<double, std::milli>(b - a).count();
                t_ws = std::chrono::duration<double, std::milli>(d - b).count();
            } catch (...) {
                c.s = s_main;
                throw;
            }
            c.s = s_main;
        }
        HIPCHK(hipEventRecord(ev.e[1], s2));
        HIPCHK(hipStreamWaitEvent(c.s, ev.e[1], 0));
        if (split) precond_finish(&c);
        auto t2b = std::chrono::steady_clock::now();
        ph[2] += std::chrono::duration<double, std::milli>(t2b - t1).count();
        int it = 0, ie = 0, used = 0;
        // without a model dxs was just zeroed: no host-side check of the guess
        gmres_dev(&c, db.p, dxs, tol, 10, 10, &it, err, &ie, model != nullptr, &used,
                  model == nullptr);
        auto t3 = std::chrono::steady_clock::now();
        energy_dev(&c, nq, qpos, q, eps_p, dxs, soleng, couleng, dev);
        auto t4 = std::chrono::steady_clock::now();
        if (x_out && !dev) {
            HIPCHK(hipMemcpyAsync(x_out, dxs, sizeof(double) * N, hipMemcpyDeviceToHost, c.s));
        }
        HIPCHK(hipStreamSynchronize(c.s));
        *iter = it;
        *ierr = ie;
        if (ws_used) *ws_used = used;
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        g_times[0] = ph[0]; g_times[1] = ph[1]; g_times[2] = ph[2];
        g_times[3] = t_rhs; g_times[4] = ms(t2b, t3); g_times[5] = ms(t3, t4);
        g_times[6] = ms(t0, t4); g_times[7] = t_ws;
    } catch (...) {
        StreamPool::release(c.aux.s);
        c.aux.s = nullptr;
        StreamPool::release(s2);
        hipStream_t s = c.s;
        c.s = nullptr;
        StreamPool::release(s);
        throw;
    }
    // back to the pool in the reverse order of their taking, so the next
    // solve gets each stream in the same role
    StreamPool::release(c.aux.s);
    c.aux.s = nullptr;
    StreamPool::release(s2);
    hipStream_t s = c.s;
    c.s = nullptr;
    StreamPool::release(s);
}

}  // namespace

extern "C" {

int tabi_solve_ws(const tabi_params *prm, int32_t n, const double *xyz, const double *nrm,
                  const double *area, int32_t nq, const double *qpos, const double *q,
                  double eps_p, double tol, tabi_ws *model, double *x_out,
                  int32_t *orderarr_out, int32_t *iter, double *err, int32_t *ierr,
                  double *soleng, double *couleng, int32_t *ws_used)
{
    return guarded([&] {
        solve_core(prm, n, xyz, nrm, area, nq, qpos, q, eps_p, tol, model, x_out, orderarr_out,
                   iter, err, ierr, soleng, couleng, ws_used, false);
    });
}

int tabi_solve_dev(const tabi_params *prm, int32_t n, const double *d_xyz, const double *d_nrm,
                   const double *d_area, int32_t nq, const double *d_qpos, const double *d_q,
                   double eps_p, double tol, tabi_ws *model, double *d_x_out,
                   int32_t *orderarr_out, int32_t *iter, double *err, int32_t *ierr,
                   double *soleng, double *couleng, int32_t *ws_used)
{
    return guarded([&] {
        solve_core(prm, n, d_xyz, d_nrm, d_area, nq, d_qpos, d_q, eps_p, tol, model, d_x_out,
                   orderarr_out, iter, err, ierr, soleng, couleng, ws_used, true);
    });
}

int64_t tabi_ws_param_count(void) { return tabi::ws::kNP; }

int tabi_ws_create(const float *params, int64_t count, tabi_ws **out)
{
    return guarded([&] {
        if (!out) throw Error("null output handle");
        *out = nullptr;
        if (!params) throw Error("null parameters");
        if (count != tabi::ws::kNP)
            throw Error("warm-start model needs " + std::to_string(tabi::ws::kNP) + " parameters, got " +
                        std::to_string(count));
        for (int k = 0; k < tabi::ws::kF; ++k)
            if (!(params[tabi::ws::oSig + k] > 0.0f)) throw Error("feature scales must be positive");
        for (int k = 0; k < tabi::ws::kO; ++k)
            if (!(params[tabi::ws::oTs + k] > 0.0f)) throw Error("output scales must be positive");
        auto *h = new tabi_ws;
        try {
            h->m = tabi::ws_create(params);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void tabi_ws_destroy(tabi_ws *m)
{
    if (!m) return;
    tabi::ws_destroy(m->m);
    delete m;
}

int tabi_ws_get_params(tabi_ws *m, float *out, int64_t count)
{
    return guarded([&] {
        if (!m || !m->m) throw Error("null warm-start model");
        if (count != tabi::ws::kNP) throw Error("parameter count mismatch");
        tabi::ws_get_params(m->m, out);
    });
}

int tabi_ctx_ws_features(tabi_ctx *c, int32_t nq, const double *qpos, const double *q,
                         double eps_p, float *feat_out, double *scale_out)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (nq < 1) throw Error("need at least one charge");
        DBuf<double> db;
        DBuf<float> feat;
        db.alloc((size_t)2 * c->n);
        feat.alloc((size_t)c->n * tabi::ws::kF);
        rhs_dev(c, nq, qpos, q, eps_p, db.p);
        const double scale = ws_features_dev(c, nq, qpos, q, eps_p, db.p, feat.p);
        HIPCHK(hipMemcpyAsync(feat_out, feat.p, sizeof(float) * c->n * tabi::ws::kF,
                              hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        if (scale_out) *scale_out = scale;
    });
}

int tabi_ws_predict(tabi_ws *m, int64_t n, const float *feat, double scale, double *x0_out)
{
    return guarded([&] {
        if (!m || !m->m) throw Error("null warm-start model");
        if (n < 0) throw Error("negative face count");
        if (!(scale > 0.0)) throw Error("scale must be positive");
        tabi::ws_predict(m->m, n, feat, scale, x0_out);
    });
}

int tabi_ws_grad(tabi_ws *m, int64_t ns, const float *feat, const float *target,
                 float *grad_out, double *loss_out)
{
    return guarded([&] {
        if (!m || !m->m) throw Error("null warm-start model");
        if (ns < 1) throw Error("need at least one sample");
        tabi::ws_grad(m->m, ns, feat, target, grad_out, loss_out);
    });
}

int tabi_ws_train(tabi_ws *m, int64_t ns, const float *feat, const float *target,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss)
{
    return guarded([&] {
        if (!m || !m->m) throw Error("null warm-start model");
        if (batch < 1 || batch > 4096 * 32) throw Error("batch must be 1..131072");
        tabi::ws_train(m->m, ns, feat, target, epochs, batch, lr, seed, epoch_loss);
    });
}

int tabi_ctx_gmres_ws(tabi_ctx *c, tabi_ws *model, int32_t nq, const double *qpos,
                      const double *q, double eps_p, const double *b, double *x, double tol,
                      int32_t maxl, int32_t nrmax, int32_t *iter, double *err, int32_t *ierr,
                      int32_t *ws_used)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (!model || !model->m) throw Error("null warm-start model");
        const int64_t N = 2 * (int64_t)c->n;
        DBuf<double> db, dxs;
        DBuf<float> feat;
        db.upload(b, N, c->s);
        dxs.alloc(N);
        feat.alloc((size_t)c->n * tabi::ws::kF);
        if (nq < 1) throw Error("need at least one charge");
        const double scale = ws_features_dev(c, nq, qpos, q, eps_p, db.p, feat.p);
        tabi::ws_predict_dev(model->m, c->n, feat.p, scale, dxs.p, c->s);
        HIPCHK(hipStreamSynchronize(c->s));
        int it = 0, ie = 0, used = 0;
        gmres_dev(c, db.p, dxs.p, tol, maxl, nrmax, &it, err, &ie, true, &used);
        HIPCHK(hipMemcpyAsync(x, dxs.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        *iter = it;
        *ierr = ie;
        if (ws_used) *ws_used = used;
    });
}

int tabi_ctx_profile(tabi_ctx *c, int32_t enable)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        if (enable && !c->ev[0]) {
            HIPCHK(hipEventCreate(&c->ev[0]));
            HIPCHK(hipEventCreate(&c->ev[1]));
        }
        c->prof = enable != 0;
        for (int i = 0; i < 8; ++i) { c->kt_ms[i] = 0.0; c->kt_n[i] = 0; }
    });
}

int tabi_ctx_kernel_times(tabi_ctx *c, double *out_ms, int64_t *out_launches)
{
    return guarded([&] {
        if (!c) throw Error("null context");
        PoolScope ps_(c->s);
        for (int i = 0; i < 8; ++i) { out_ms[i] = c->kt_ms[i]; out_launches[i] = c->kt_n[i]; }
    });
}

int tabi_last_solve_times(double *out8)
{
    for (int i = 0; i < 8; ++i) out8[i] = g_times[i];
    return 0;
}

}  // extern "C"
