// Learned initial guess for GMRES: a per-face MLP (12 features -> 3 x 128 GELU
// -> 2 outputs, the scaled surface potential phi and its normal derivative)
// with BF16 weights and activations and FP32 accumulation on the matrix cores.
// There is no counterpart in the reference; the guess enters where the
// reference sets xvct = 0 before DGMRES (src/main.f90:115), in tree order.
//
// Kernels
//   k_ws_feat    per-face features from the tree-order geometry and the charges
//   k_ws_pack    FP32 master weights -> padded BF16 weight images
//   k_ws_mlp<T>  forward (T=false: writes the guess) or forward + backward
//                (T=true: writes one gradient partial per workgroup)
//   k_ws_reduce  ordered sum of the partials; k_ws_adam the optimiser step
//
// MFMA use (v_mfma_f32_16x16x16_bf16): every product is laid out neurons x
// faces, D[n][f] = sum_k W[n][k] act[k][f], so a lane's output registers
// (rows 4*(lane/16)+r, column lane%16) hold exactly what the next layer's
// B operand needs (rows 4*(lane/16)+0..3 of the same column).  Activations and
// deltas pass through LDS only where every wave needs every tile (the next
// layer, and the weight gradients, whose contraction runs over faces).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "tabi_internal.h"

namespace tabi {
namespace ws {
namespace {

#define WCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kChunk = 32;             // faces per workgroup pass
constexpr int kThreads = 256;          // 4 waves
constexpr int LW = 136;                // padded LDS row stride, 128-wide rows
constexpr int LK = 20;                 // padded LDS row stride, 16-wide rows
constexpr int kMaxChunks = 4096;       // gradient partials (chunks) per step
// BF16 weight image (uint16 offsets): W1 [128][16], W2 [128][128], W3 [128][128],
// W4 [16][128] (rows 2..15 and columns 12..15 of W1 are zero)
constexpr int iW1 = 0, iW2 = iW1 + kH * 16, iW3 = iW2 + kH * kH, iW4 = iW3 + kH * kH,
              kImg = iW4 + 16 * kH;

struct Lds {
    uint16_t w1[kH * LK], w2[kH * LW], w3[kH * LW], w4[16 * LW];
    uint16_t a0[kChunk * LK], a1[kChunk * LW], a2[kChunk * LW], a3[kChunk * LW];
    uint16_t d4[kChunk * LK], da[kChunk * LW], db[kChunk * LW];
};

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ s4 ld4(const uint16_t *p)
{
    return *reinterpret_cast<const s4 *>(p);
}

__device__ __forceinline__ s4 ld4s(const uint16_t *p, int stride)
{
    s4 v;
    v.x = (short)p[0];
    v.y = (short)p[stride];
    v.z = (short)p[2 * stride];
    v.w = (short)p[3 * stride];
    return v;
}

__device__ __forceinline__ void st4(uint16_t *p, float a, float b, float c, float d)
{
    s4 v;
    v.x = (short)f2bf(a); v.y = (short)f2bf(b); v.z = (short)f2bf(c); v.w = (short)f2bf(d);
    *reinterpret_cast<s4 *>(p) = v;
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// tanh(u) = 1 - 2 / (exp(2u) + 1) with the hardware exponential and
// reciprocal (absolute error ~1e-7; exp overflow gives 1, underflow -1)
__device__ __forceinline__ float tanh_fast(float u)
{
    const float e = __expf(2.0f * u);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// GELU, tanh form, and its derivative
__device__ __forceinline__ float gelu(float z)
{
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return 0.5f * z * (1.0f + tanh_fast(u));
}

__device__ __forceinline__ float gelu_d(float z)
{
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    const float t = tanh_fast(u);
    return 0.5f * (1.0f + t) +
           0.5f * z * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 0.134145f * z * z);
}

}  // namespace

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// Per face i (tree order), written unnormalised to feat[i][0..11], with
// s = 1/rms(b_1..b_n) (so features and the scaled solution do not depend on
// the overall charge scale), u = 1/(4 pi eps_p), r the centroid-charge
// distance, cos the cosine between the normal and the centroid-to-charge
// direction, and Gaussian windows w1 = exp(-r^2/R1^2), w2 = exp(-r^2/R2^2):
//   0  s * b_i                 Coulomb potential of all charges (right-hand side)
//   1  s * b_{n+i}             its normal derivative (right-hand side)
//   2  s*u*sum q w1/r          potential of the charges near the face (R1 = 5 A)
//   3  s*u*sum q w1 cos/r^2    their normal field
//   4  s*u*sum q w2/r          the same within R2 = 10 A
//   5  s*u*sum q w2 cos/r^2
//   6  area                    7  curvature: Gaussian-weighted mean of
//                                 (n_j-n_i).(c_j-c_i)/|c_j-c_i|^2 over faces j
//                                 of the same leaf (exactly 1/R on a sphere)
//   8-9 distances to the nearest and second-nearest charge
//   10 normal . (centroid - nearest charge)/distance
//   11 sum w2 (how buried the face is: a smooth count of charges within R2)
constexpr float kCurvH = 1.0f;      // curvature kernel width (Angstrom)
constexpr float kR1 = 5.0f, kR2 = 10.0f;

// The charges are visited in tiles of kQTile that are compact in space (sorted
// by a Morton code of their position on a 1 A grid), each with a bounding
// sphere.  A wave skips a tile when, for every one of its faces, the sphere
// lies beyond the widest window (3 R2) and beyond the face's second-nearest
// charge so far: no charge of the tile could then change any sum or the
// nearest-charge features.  Each wave starts at the tile of its own position,
// so the nearest charges come first.
constexpr int kQTile = 32;
constexpr float kQSkip = 9.0f * 100.0f;     // (3 R2)^2, R2 = 10 A
constexpr float kMortonOff = 512.0f;        // grid origin -512 A, 1 A cells, 10 bits per axis

__device__ __forceinline__ uint32_t morton10(float x, float y, float z)
{
    auto spread = [](uint32_t v) {
        v &= 0x3ffu;
        v = (v | (v << 16)) & 0x030000ffu;
        v = (v | (v << 8)) & 0x0300f00fu;
        v = (v | (v << 4)) & 0x030c30c3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    auto cell = [](float c) {
        const float g = floorf(c + kMortonOff);
        return (uint32_t)fminf(fmaxf(g, 0.0f), 1023.0f);
    };
    return spread(cell(x)) | (spread(cell(y)) << 1) | (spread(cell(z)) << 2);
}

__global__ void k_ws_qkey(int nq, const double *__restrict__ qpos, uint32_t *__restrict__ key,
                          int32_t *__restrict__ idx)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nq) return;
    key[a] = morton10((float)qpos[3 * a], (float)qpos[3 * a + 1], (float)qpos[3 * a + 2]);
    idx[a] = a;
}

// charge records {x, y, z, q} in FP32, in Morton order
__global__ void k_ws_qrec(int nq, const double *__restrict__ qpos, const double *__restrict__ qch,
                          const int32_t *__restrict__ idx, float4 *__restrict__ rec)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const int a = idx[t];
    rec[t] = make_float4((float)qpos[3 * a], (float)qpos[3 * a + 1], (float)qpos[3 * a + 2], (float)qch[a]);
}

// bounding sphere of each tile {centre, radius}: the centre of the tile's box,
// the radius its largest distance to a charge (plus a margin for rounding)
__global__ void k_ws_qtile(int nq, const float4 *__restrict__ rec, float4 *__restrict__ tile)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int a0 = t * kQTile, a1 = min(nq, a0 + kQTile);
    if (a0 >= nq) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int a = a0; a < a1; ++a) {
        const float4 c = rec[a];
        lo[0] = fminf(lo[0], c.x); hi[0] = fmaxf(hi[0], c.x);
        lo[1] = fminf(lo[1], c.y); hi[1] = fmaxf(hi[1], c.y);
        lo[2] = fminf(lo[2], c.z); hi[2] = fmaxf(hi[2], c.z);
    }
    const float mx = 0.5f * (lo[0] + hi[0]), my = 0.5f * (lo[1] + hi[1]), mz = 0.5f * (lo[2] + hi[2]);
    float r2 = 0.0f;
    for (int a = a0; a < a1; ++a) {
        const float4 c = rec[a];
        const float dx = c.x - mx, dy = c.y - my, dz = c.z - mz;
        r2 = fmaxf(r2, dx * dx + dy * dy + dz * dz);
    }
    tile[t] = make_float4(mx, my, mz, sqrtf(r2) * 1.0001f + 1.0e-3f);
}

// Each wave stages the records of the tile it visits in its own LDS slots
// (one coalesced load), then every lane reads them as broadcasts.
__global__ void __launch_bounds__(256) k_ws_feat(int n, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz, const double *__restrict__ tnx,
    const double *__restrict__ tny, const double *__restrict__ tnz,
    const double *__restrict__ tarea, const int32_t *__restrict__ leaf_of,
    const int32_t *__restrict__ leaf_beg, const int32_t *__restrict__ leaf_m2, int nq,
    const float4 *__restrict__ qrec, const float4 *__restrict__ qtile,
    const uint32_t *__restrict__ qkey, double eps_p,
    const double *__restrict__ b, double s, float *__restrict__ feat)
{
    __shared__ float4 s_rec[256 / 64][kQTile];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const bool valid = i < n;
    const int ii = valid ? i : n - 1;
    const float cx = (float)tx[ii], cy = (float)ty[ii], cz = (float)tz[ii];
    const float nx = (float)tnx[ii], ny = (float)tny[ii], nz = (float)tnz[ii];
    static_assert(kR2 == 2.0f * kR1, "the narrow window is derived from the wide one");
    const float i2 = 1.0f / (kR2 * kR2);
    float d1 = INFINITY, d2 = INFINITY;          // squared distances, ascending
    float ax = 0.f, ay = 0.f, az = 0.f;          // nearest charge
    float p1 = 0.f, f1 = 0.f, p2 = 0.f, f2 = 0.f, bur = 0.f;
    static_assert(kQSkip == 9.0f * kR2 * kR2, "the skip bound is the wide window's cutoff");
    const int ntile = (nq + kQTile - 1) / kQTile;
    // the wave's first tile: the first whose charges' Morton codes reach the
    // code of the wave's first face (a binary search in the sorted codes)
    int t0 = 0;
    if (ntile > 0) {
        const uint32_t key = morton10(__shfl(cx, 0), __shfl(cy, 0), __shfl(cz, 0));
        int lo = 0, hi = nq;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (qkey[mid] < key) lo = mid + 1; else hi = mid;
        }
        t0 = min(lo / kQTile, ntile - 1);
    }
    for (int tt = 0; tt < ntile; ++tt) {
        int t = t0 + tt;
        if (t >= ntile) t -= ntile;
        const float4 sp = qtile[t];
        const float ex = sp.x - cx, ey = sp.y - cy, ez = sp.z - cz;
        const float dt = sqrtf(ex * ex + ey * ey + ez * ez) - sp.w;
        const bool far = valid && dt > 0.0f && dt * dt > fmaxf(kQSkip, d2);
        if (__all(far || !valid)) continue;
        const int a0 = t * kQTile, na = min(nq - a0, kQTile);
        __builtin_amdgcn_wave_barrier();          // the previous tile's reads are done
        if (ln < na) s_rec[wv][ln] = qrec[a0 + ln];
        __builtin_amdgcn_wave_barrier();
#pragma unroll 4
        for (int k = 0; k < na; ++k) {
            const float4 c = s_rec[wv][k];
            const float dx = c.x - cx, dy = c.y - cy, dz = c.z - cz;
            const float r2 = dx * dx + dy * dy + dz * dz;
            // the two nearest (squared) distances, first nearest charge kept
            const bool closer = r2 < d1;
            d2 = fminf(d2, fmaxf(d1, r2));
            d1 = fminf(d1, r2);
            ax = closer ? c.x : ax;
            ay = closer ? c.y : ay;
            az = closer ? c.z : az;
            if (r2 < 9.0f * kR2 * kR2 && r2 > 0.f) {
                const float ir = __builtin_amdgcn_rsqf(r2);
                // kR1 = kR2 / 2: exp(-r2/kR1^2) = exp(-r2/kR2^2)^4
                const float w2 = __expf(-r2 * i2), w2s = w2 * w2, w1 = w2s * w2s;
                const float pot = c.w * ir;
                const float fld = pot * (nx * dx + ny * dy + nz * dz) * ir * ir;
                p1 += pot * w1; f1 += fld * w1;
                p2 += pot * w2; f2 += fld * w2;
                bur += w2;
            }
        }
    }
    if (!valid) return;
    // curvature over the leaf
    const double dnx = tnx[i], dny = tny[i], dnz = tnz[i];
    const int l = leaf_of[i];
    const int j0 = leaf_beg[l], j1 = j0 + leaf_m2[l] / 2;
    const double xi = tx[i], yi = ty[i], zi = tz[i];
    const double ih2 = 1.0 / ((double)kCurvH * kCurvH);
    double num = 0.0, den = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double dx = tx[j] - xi, dy = ty[j] - yi, dz = tz[j] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (j == i || r2 <= 0.0 || r2 * ih2 > 9.0) continue;
        const double w = exp(-r2 * ih2);
        num += w * ((tnx[j] - dnx) * dx + (tny[j] - dny) * dy + (tnz[j] - dnz) * dz) / r2;
        den += w;
    }
    const double su = s * 0.25 / (3.141592653589793 * eps_p);
    float *f = feat + (size_t)12 * i;
    f[0] = (float)(s * b[i]);
    f[1] = (float)(s * b[n + i]);
    f[2] = (float)(su * p1);
    f[3] = (float)(su * f1);
    f[4] = (float)(su * p2);
    f[5] = (float)(su * f2);
    f[6] = (float)tarea[i];
    f[7] = den > 0.0 ? (float)(num / den) : 0.0f;
    const float r1 = d1 < INFINITY ? sqrtf(d1) : 0.0f;
    f[8] = r1;
    f[9] = d2 < INFINITY ? sqrtf(d2) : r1;
    f[10] = r1 > 0.f ? (nx * (cx - ax) + ny * (cy - ay) + nz * (cz - az)) / r1 : 0.0f;
    f[11] = bur;
}

// ---------------------------------------------------------------------------
// MLP
// ---------------------------------------------------------------------------
__global__ void k_ws_pack(const float *__restrict__ P, uint16_t *__restrict__ img)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kImg) return;
    float v = 0.0f;
    if (e < iW2) {
        const int nrow = e / 16, k = e % 16;
        v = k < kF ? P[oW1 + nrow * kF + k] : 0.0f;
    } else if (e < iW3) {
        v = P[oW2 + (e - iW2)];
    } else if (e < iW4) {
        v = P[oW3 + (e - iW3)];
    } else {
        const int o = (e - iW4) / kH, k = (e - iW4) % kH;
        v = o < kO ? P[oW4 + o * kH + k] : 0.0f;
    }
    img[e] = f2bf(v);
}

// Forward for this wave's four output tiles (neuron tiles nt0, nt0+1; face
// tiles 0, 1) of a 128-wide layer.
template <int KB>
__device__ __forceinline__ void fwd_tiles(const uint16_t *W, int ldw, const uint16_t *act, int lda,
                                          int nt0, int lane, f4 (&acc)[2][2])
{
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        s4 a[2], bb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = ld4(W + (16 * (nt0 + i) + r16) * ldw + 16 * kb + 4 * g);
#pragma unroll
        for (int j = 0; j < 2; ++j) bb[j] = ld4(act + (16 * j + r16) * lda + 16 * kb + 4 * g);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], bb[j], acc[i][j]);
    }
}

// z = acc + bias, act_out = bf16(gelu(z)); z kept for the backward pass
__device__ __forceinline__ void bias_gelu_store(const float *bias, int nt0, int lane,
                                                f4 (&acc)[2][2], uint16_t *out)
{
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int n0 = 16 * (nt0 + i) + 4 * g;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            f4 &z = acc[i][j];
            z.x += bias[n0]; z.y += bias[n0 + 1]; z.z += bias[n0 + 2]; z.w += bias[n0 + 3];
            st4(out + (16 * j + r16) * LW + n0, gelu(z.x), gelu(z.y), gelu(z.z), gelu(z.w));
        }
    }
}

// delta_in[n][f] = (sum_m W[m][n] delta[f][m]) * gelu'(z[n][f]) for the wave's
// tiles; written as BF16 to dout and returned (FP32) for the bias gradient
template <int KB>
__device__ __forceinline__ void bwd_delta(const uint16_t *W, int ldw, const uint16_t *dlt, int ldd,
                                          int nt0, int lane, const f4 (&z)[2][2], f4 (&dl)[2][2],
                                          uint16_t *dout)
{
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) dl[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int kb = 0; kb < KB; ++kb) {
        s4 a[2], bb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = ld4s(W + (16 * kb + 4 * g) * ldw + 16 * (nt0 + i) + r16, ldw);
#pragma unroll
        for (int j = 0; j < 2; ++j) bb[j] = ld4(dlt + (16 * j + r16) * ldd + 16 * kb + 4 * g);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) dl[i][j] = mfma(a[i], bb[j], dl[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            f4 &d = dl[i][j];
            const f4 &zz = z[i][j];
            d.x *= gelu_d(zz.x); d.y *= gelu_d(zz.y); d.z *= gelu_d(zz.z); d.w *= gelu_d(zz.w);
            st4(dout + (16 * j + r16) * LW + 16 * (nt0 + i) + 4 * g, d.x, d.y, d.z, d.w);
        }
}

// dW[m][n] = sum_f delta[f][m] act[f][n] over the chunk's 32 faces, rows m in
// tiles mt0, mt0+1, NT column tiles; stored to dst[m*ldo + n] for n < ldo
template <int NT, int ldo>
__device__ __forceinline__ void wgrad_store(const uint16_t *dlt, int ldd, const uint16_t *act,
                                            int lda, int mt0, int lane, float *dst)
{
    const int r16 = lane & 15, g = lane >> 4;
    s4 a[2][2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[kb][i] = ld4s(dlt + (16 * kb + 4 * g) * ldd + 16 * (mt0 + i) + r16, ldd);
#pragma unroll 2
    for (int nt = 0; nt < NT; ++nt) {
        f4 acc[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const s4 bb = ld4s(act + (16 * kb + 4 * g) * lda + 16 * nt + r16, lda);
#pragma unroll
            for (int i = 0; i < 2; ++i) acc[i] = mfma(a[kb][i], bb, acc[i]);
        }
        const int n = 16 * nt + r16;
        if (n < ldo) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int m0 = 16 * (mt0 + i) + 4 * g;
                dst[(m0 + 0) * ldo + n] = acc[i].x;
                dst[(m0 + 1) * ldo + n] = acc[i].y;
                dst[(m0 + 2) * ldo + n] = acc[i].z;
                dst[(m0 + 3) * ldo + n] = acc[i].w;
            }
        }
    }
}

// per-lane sums over the two face tiles of the wave's delta tiles
__device__ __forceinline__ void bias_sum(const f4 (&dl)[2][2], float (&gb)[2][4])
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        gb[i][0] = dl[i][0].x + dl[i][1].x;
        gb[i][1] = dl[i][0].y + dl[i][1].y;
        gb[i][2] = dl[i][0].z + dl[i][1].z;
        gb[i][3] = dl[i][0].w + dl[i][1].w;
    }
}

// bias gradient: sum over the 16 face lanes of each lane group
__device__ __forceinline__ void write_bias(float *dst, float (&gb)[2][4], int nt0, int lane)
{
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = gb[i][r];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
            if (r16 == 0) dst[16 * (nt0 + i) + 4 * g + r] = v;
        }
}

// Stage the BF16 weight image into LDS.
__device__ __forceinline__ void stage_weights(Lds &L, const uint16_t *__restrict__ img)
{
    for (int e = threadIdx.x; e < kImg; e += kThreads) {
        const uint16_t v = img[e];
        if (e < iW2) L.w1[(e / 16) * LK + e % 16] = v;
        else if (e < iW3) L.w2[((e - iW2) / kH) * LW + (e - iW2) % kH] = v;
        else if (e < iW4) L.w3[((e - iW3) / kH) * LW + (e - iW3) % kH] = v;
        else L.w4[((e - iW4) / kH) * LW + (e - iW4) % kH] = v;
    }
}

// Inference (TRAIN=false): x0[f] = out0 * sc0, x0[n+f] = out1 * sc1; the
// workgroups stride over the 32-face chunks.
// Training (TRAIN=true): one chunk per workgroup (grid = number of chunks) of
// samples idx[first .. first+ns-1] (or first.. when idx is null), loss
// (1/2B) sum (out - t/ts)^2; the workgroup's gradient partial goes to row
// blockIdx.x of `partial` in the packed parameter layout, its loss sum to slot kNP.
template <bool TRAIN>
__global__ void __launch_bounds__(kThreads) k_ws_mlp(int ns, const float *__restrict__ feat,
    const int32_t *__restrict__ idx, int64_t first, const float *__restrict__ P,
    const uint16_t *__restrict__ img, double *__restrict__ x0, int64_t nface, double sc0,
    double sc1, const float *__restrict__ target, float inv_batch, float *__restrict__ partial)
{
    __shared__ Lds L;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r16 = lane & 15, g = lane >> 4;
    const int nt0 = 2 * w;
    stage_weights(L, img);
    const float *b1 = P + oB1, *b2 = P + oB2, *b3 = P + oB3, *b4 = P + oB4;
    const float *mu = P + oMu, *sig = P + oSig, *ts = P + oTs;
    const int nchunk = (ns + kChunk - 1) / kChunk;
    for (int ch = blockIdx.x; ch < nchunk; ch += TRAIN ? nchunk : gridDim.x) {
        const int base = ch * kChunk;
        __syncthreads();  // weights staged / previous chunk's readers are done
        // features -> a0 (normalised, BF16); zero rows beyond ns, zero columns 12..15
        for (int e = threadIdx.x; e < kChunk * 16; e += kThreads) {
            const int f = e / 16, k = e % 16;
            float v = 0.0f;
            if (base + f < ns && k < kF) {
                const int64_t sidx = idx ? (int64_t)idx[first + base + f] : first + base + f;
                v = (feat[sidx * kF + k] - mu[k]) / sig[k];
            }
            L.a0[f * LK + k] = f2bf(v);
        }
        __syncthreads();
        f4 z1[2][2], z2[2][2], z3[2][2];
        fwd_tiles<1>(L.w1, LK, L.a0, LK, nt0, lane, z1);
        bias_gelu_store(b1, nt0, lane, z1, L.a1);
        __syncthreads();
        fwd_tiles<8>(L.w2, LW, L.a1, LW, nt0, lane, z2);
        bias_gelu_store(b2, nt0, lane, z2, L.a2);
        __syncthreads();
        fwd_tiles<8>(L.w3, LW, L.a2, LW, nt0, lane, z3);
        bias_gelu_store(b3, nt0, lane, z3, L.a3);
        __syncthreads();
        // output layer: wave 0, both face tiles; lane holds outputs o = 4g+r of
        // face 16j + r16
        float gb4_0 = 0.f, gb4_1 = 0.f, loss = 0.f;
        if (w == 0) {
            f4 o[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
                const s4 a = ld4(L.w4 + r16 * LW + 16 * kb + 4 * g);
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    o[j] = mfma(a, ld4(L.a3 + (16 * j + r16) * LW + 16 * kb + 4 * g), o[j]);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int f = base + 16 * j + r16;
                const float o0 = o[j].x + b4[0], o1 = o[j].y + b4[1];
                if (!TRAIN) {
                    if (g == 0 && f < ns) {
                        x0[f] = (double)o0 * sc0;
                        x0[nface + f] = (double)o1 * sc1;
                    }
                } else {
                    float d0 = 0.f, d1 = 0.f;
                    if (g == 0 && f < ns) {
                        const int64_t sidx = idx ? (int64_t)idx[first + f] : first + f;
                        const float e0 = o0 - target[2 * sidx] / ts[0];
                        const float e1 = o1 - target[2 * sidx + 1] / ts[1];
                        loss += e0 * e0 + e1 * e1;
                        d0 = e0 * inv_batch;
                        d1 = e1 * inv_batch;
                        gb4_0 += d0;
                        gb4_1 += d1;
                    }
                    // lanes of group g hold outputs 4g..4g+3: only 0 and 1 are real
                    st4(L.d4 + (16 * j + r16) * LK + 4 * g, d0, d1, 0.f, 0.f);
                }
            }
        }
        if (!TRAIN) continue;
        float *out = partial + (size_t)blockIdx.x * (kNP + 1);
        __syncthreads();
        // layer-3 delta and the output-layer weight gradient
        f4 dl[2][2];
        float gb[2][4];
        bwd_delta<1>(L.w4, LW, L.d4, LK, nt0, lane, z3, dl, L.da);
        bias_sum(dl, gb);
        write_bias(out + oB3, gb, nt0, lane);
        {
            // dW4[o][n] = sum_f d4[f][o] a3[f][n]: rows o (one tile), columns nt0, nt0+1
            f4 gw4[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const s4 a = ld4s(L.d4 + (16 * kb + 4 * g) * LK + r16, LK);
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    gw4[i] = mfma(a, ld4s(L.a3 + (16 * kb + 4 * g) * LW + 16 * (nt0 + i) + r16, LW),
                                  gw4[i]);
            }
            // rows o = 4g + r: only o < 2 are real
            if (g == 0) {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    out[oW4 + 16 * (nt0 + i) + r16] = gw4[i].x;
                    out[oW4 + kH + 16 * (nt0 + i) + r16] = gw4[i].y;
                }
            }
        }
        __syncthreads();
        wgrad_store<8, kH>(L.da, LW, L.a2, LW, nt0, lane, out + oW3);
        bwd_delta<8>(L.w3, LW, L.da, LW, nt0, lane, z2, dl, L.db);
        bias_sum(dl, gb);
        write_bias(out + oB2, gb, nt0, lane);
        __syncthreads();
        wgrad_store<8, kH>(L.db, LW, L.a1, LW, nt0, lane, out + oW2);
        bwd_delta<8>(L.w2, LW, L.db, LW, nt0, lane, z1, dl, L.da);
        bias_sum(dl, gb);
        write_bias(out + oB1, gb, nt0, lane);
        __syncthreads();
        wgrad_store<1, kF>(L.da, LW, L.a0, LK, nt0, lane, out + oW1);
        if (w == 0) {
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                gb4_0 += __shfl_xor(gb4_0, off);
                gb4_1 += __shfl_xor(gb4_1, off);
                loss += __shfl_xor(loss, off);
            }
            if (lane == 0) {
                out[oB4] = gb4_0;
                out[oB4 + 1] = gb4_1;
                out[kNP] = loss;
            }
            if (lane < oW1) out[lane] = 0.0f;
        }
    }
}

// grad[p] = sum over workgroups (fixed order); loss_acc[0] += loss sum
__global__ void k_ws_reduce(int nwg, const float *__restrict__ partial, float *__restrict__ grad,
                            double *__restrict__ loss_acc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kNP) return;
    float s = 0.0f;
    for (int wg = 0; wg < nwg; ++wg) s += partial[(size_t)wg * (kNP + 1) + p];
    if (p < kNP) grad[p] = p < oW1 ? 0.0f : s;
    else if (loss_acc) loss_acc[0] += (double)s;
}

// Adam (bias-corrected) on the trainable parameters
__global__ void k_ws_adam(float *__restrict__ P, const float *__restrict__ grad,
                          float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                          float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x + oW1;
    if (p >= kNP) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
struct Dev {
    T *p = nullptr;
    size_t n = 0;
    Dev() = default;
    explicit Dev(size_t cnt) { alloc(cnt); }
    ~Dev() { device_free(p); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
    void alloc(size_t cnt)
    {
        if (p && cnt <= n) return;
        device_free(p);
        p = static_cast<T *>(device_alloc(std::max<size_t>(cnt, 1) * sizeof(T)));
        n = cnt;
    }
};

inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace ws

struct WsModel {
    ws::Dev<float> P, m, v, grad, partial;
    ws::Dev<uint16_t> img;
    ws::Dev<double> loss;
    hipStream_t s = nullptr;
    int64_t step = 0;
};

using namespace ws;

WsModel *ws_create(const float *params)
{
    auto *M = new WsModel;
    try {
        WCHK(hipStreamCreate(&M->s));
        M->P.alloc(kNP);
        M->m.alloc(kNP);
        M->v.alloc(kNP);
        M->grad.alloc(kNP);
        M->img.alloc(kImg);
        M->loss.alloc(1);
        WCHK(hipMemcpyAsync(M->P.p, params, sizeof(float) * kNP, hipMemcpyHostToDevice, M->s));
        WCHK(hipMemsetAsync(M->m.p, 0, sizeof(float) * kNP, M->s));
        WCHK(hipMemsetAsync(M->v.p, 0, sizeof(float) * kNP, M->s));
        k_ws_pack<<<nb(kImg, 256), 256, 0, M->s>>>(M->P.p, M->img.p);
        WCHK(hipGetLastError());
        WCHK(hipStreamSynchronize(M->s));
    } catch (...) {
        ws_destroy(M);
        throw;
    }
    return M;
}

void ws_destroy(WsModel *M)
{
    if (!M) return;
    if (M->s) {
        (void)hipStreamSynchronize(M->s);
        (void)hipStreamDestroy(M->s);
    }
    delete M;
}

void ws_get_params(WsModel *M, float *out)
{
    WCHK(hipMemcpyAsync(out, M->P.p, sizeof(float) * kNP, hipMemcpyDeviceToHost, M->s));
    WCHK(hipStreamSynchronize(M->s));
}

void ws_features(const WsFeatIn &in, float *d_feat, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (in.n <= 0) return;
    // charges sorted into compact tiles (Morton order), records and spheres
    const int nq = std::max(in.nq, 0), ntile = (nq + kQTile - 1) / kQTile;
    Dev<float4> qrec(std::max(nq, 1)), qtile(std::max(ntile, 1));
    Dev<uint32_t> key(std::max(nq, 1)), key2(std::max(nq, 1));
    Dev<int32_t> idx(std::max(nq, 1)), idx2(std::max(nq, 1));
    if (nq > 0) {
        k_ws_qkey<<<nb(nq, 256), 256, 0, s>>>(nq, in.qpos, key.p, idx.p);
        size_t tmp_bytes = 0;
        WCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key2.p, idx.p, idx2.p, nq, 0,
                                                30, s));
        Dev<char> tmp(tmp_bytes);
        WCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key2.p, idx.p, idx2.p, nq, 0,
                                                30, s));
        k_ws_qrec<<<nb(nq, 256), 256, 0, s>>>(nq, in.qpos, in.q, idx2.p, qrec.p);
        k_ws_qtile<<<nb(ntile, 64), 64, 0, s>>>(nq, qrec.p, qtile.p);
    }
    k_ws_feat<<<nb(in.n, 256), 256, 0, s>>>(in.n, in.tx, in.ty, in.tz, in.tnx, in.tny, in.tnz,
        in.tarea, in.leaf_of, in.leaf_beg, in.leaf_m2, nq, qrec.p, qtile.p, key2.p, in.eps_p, in.b,
        in.scale, d_feat);
    WCHK(hipGetLastError());
}

void ws_predict_dev(WsModel *M, int64_t n, const float *d_feat, double scale, double *d_x0,
                    void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n <= 0) return;
    if (n > INT32_MAX) throw std::runtime_error("too many faces");
    // output scales (host copy of the header)
    float hdr[oW1];
    WCHK(hipMemcpyAsync(hdr, M->P.p, sizeof(hdr), hipMemcpyDeviceToHost, M->s));
    WCHK(hipStreamSynchronize(M->s));
    const int nchunk = (int)((n + kChunk - 1) / kChunk);
    const int grid = std::min(nchunk, 1024);
    k_ws_mlp<false><<<grid, kThreads, 0, s>>>((int)n, d_feat, nullptr, 0, M->P.p, M->img.p, d_x0, n,
        (double)hdr[oTs] / scale, (double)hdr[oTs + 1] / scale, nullptr, 0.f, nullptr);
    WCHK(hipGetLastError());
}

double ws_grad_dev(WsModel *M, int64_t ns, const float *d_feat, const int32_t *d_idx, int64_t first,
                   const float *d_target, float inv_batch, bool sync_loss)
{
    if (ns > INT32_MAX) throw std::runtime_error("batch too large");
    const int grid = (int)((ns + kChunk - 1) / kChunk);
    if (grid > kMaxChunks) throw std::runtime_error("batch larger than " + std::to_string(kMaxChunks * kChunk) + " samples");
    M->partial.alloc((size_t)grid * (kNP + 1));
    k_ws_mlp<true><<<grid, kThreads, 0, M->s>>>((int)ns, d_feat, d_idx, first, M->P.p, M->img.p,
        nullptr, 0, 0.0, 0.0, d_target, inv_batch, M->partial.p);
    k_ws_reduce<<<nb(kNP + 1, 256), 256, 0, M->s>>>(grid, M->partial.p, M->grad.p, M->loss.p);
    WCHK(hipGetLastError());
    if (!sync_loss) return 0.0;
    double l = 0.0;
    WCHK(hipMemcpyAsync(&l, M->loss.p, sizeof(double), hipMemcpyDeviceToHost, M->s));
    WCHK(hipStreamSynchronize(M->s));
    return l;
}

void ws_grad(WsModel *M, int64_t ns, const float *feat, const float *target, float *grad_out,
             double *loss_out)
{
    Dev<float> df((size_t)ns * kF), dt((size_t)ns * 2);
    WCHK(hipMemcpyAsync(df.p, feat, sizeof(float) * ns * kF, hipMemcpyHostToDevice, M->s));
    WCHK(hipMemcpyAsync(dt.p, target, sizeof(float) * ns * 2, hipMemcpyHostToDevice, M->s));
    WCHK(hipMemsetAsync(M->loss.p, 0, sizeof(double), M->s));
    const double l = ws_grad_dev(M, ns, df.p, nullptr, 0, dt.p, 1.0f / (float)ns, true);
    WCHK(hipMemcpyAsync(grad_out, M->grad.p, sizeof(float) * kNP, hipMemcpyDeviceToHost, M->s));
    WCHK(hipStreamSynchronize(M->s));
    if (loss_out) *loss_out = l / (2.0 * (double)ns);
}

void ws_predict(WsModel *M, int64_t n, const float *feat, double scale, double *x0)
{
    Dev<float> df((size_t)n * kF);
    Dev<double> dx((size_t)2 * n);
    WCHK(hipMemcpyAsync(df.p, feat, sizeof(float) * n * kF, hipMemcpyHostToDevice, M->s));
    ws_predict_dev(M, n, df.p, scale, dx.p, M->s);
    WCHK(hipMemcpyAsync(x0, dx.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, M->s));
    WCHK(hipStreamSynchronize(M->s));
}

void ws_train(WsModel *M, int64_t ns, const float *feat, const float *target, int epochs,
              int batch, float lr, uint64_t seed, float *epoch_loss)
{
    if (ns < 1 || epochs < 1 || batch < 1) throw std::runtime_error("empty training set");
    if (ns > INT32_MAX) throw std::runtime_error("training set too large");
    Dev<float> df((size_t)ns * kF), dt((size_t)ns * 2);
    Dev<int32_t> didx(ns);
    WCHK(hipMemcpyAsync(df.p, feat, sizeof(float) * ns * kF, hipMemcpyHostToDevice, M->s));
    WCHK(hipMemcpyAsync(dt.p, target, sizeof(float) * ns * 2, hipMemcpyHostToDevice, M->s));
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(ns);
    const int64_t steps_per_epoch = (ns + batch - 1) / batch;
    const int64_t total = steps_per_epoch * epochs;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int64_t t = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int64_t i = 0; i < ns; ++i) perm[i] = (int32_t)i;
        std::shuffle(perm.begin(), perm.end(), rng);
        WCHK(hipMemcpyAsync(didx.p, perm.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, M->s));
        WCHK(hipMemsetAsync(M->loss.p, 0, sizeof(double), M->s));
        for (int64_t b0 = 0; b0 < ns; b0 += batch) {
            const int64_t nb_ = std::min<int64_t>(batch, ns - b0);
            ws_grad_dev(M, nb_, df.p, didx.p, b0, dt.p, 1.0f / (float)nb_, false);
            ++t;
            ++M->step;
            // cosine decay to lr/20
            const double frac = (double)(t - 1) / (double)std::max<int64_t>(total - 1, 1);
            const float lrt = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
            const float c1 = 1.0f - std::pow(b1, (float)M->step);
            const float c2 = 1.0f - std::pow(b2, (float)M->step);
            k_ws_adam<<<nb(kNP - oW1, 256), 256, 0, M->s>>>(M->P.p, M->grad.p, M->m.p, M->v.p, lrt,
                                                           b1, b2, eps, c1, c2);
            k_ws_pack<<<nb(kImg, 256), 256, 0, M->s>>>(M->P.p, M->img.p);
            WCHK(hipGetLastError());
        }
        double l = 0.0;
        WCHK(hipMemcpyAsync(&l, M->loss.p, sizeof(double), hipMemcpyDeviceToHost, M->s));
        WCHK(hipStreamSynchronize(M->s));
        if (epoch_loss) epoch_loss[e] = (float)(l / (2.0 * (double)ns));
    }
}

}  // namespace tabi
