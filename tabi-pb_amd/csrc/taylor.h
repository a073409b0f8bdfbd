// Cartesian Taylor coefficients of the screened-Coulomb kernel for gfx950.
//
// taylor<L, KAPPA>() fills a[] with a_k = (1/k!) d^k/dy^k [exp(-kappa|x-y|)/|x-y|]
// at y = c for every multi-index k of total degree <= L, where (dx,dy,dz) = x - c.
// The coefficients are produced by the same recurrence relations, assigned to
// the same classes of multi-indices, as the reference (COMP_TCOEFF,
// src/treecode3d_pb.f:684-855): a scalar recurrence along each axis, then
// indices with a 1 and a 0, with one 0, with two 1s, with one 1, and all >= 2;
// each class uses the auxiliary coefficients b_k of exp(-kappa|x-y|).  Keeping
// the reference's assignment of relations to indices keeps the rounding
// behaviour of each coefficient close to the reference's.  Unlike the reference
// (which always runs to p+3) only degrees <= L = p+2 are formed: the far-field
// contraction never reads higher ones and no lower coefficient depends on them.
//
// With KAPPA == false the kernel is 1/|x-y| and every b_k other than b_000 = 1
// vanishes, so those terms are dropped.
//
// In FP32 the reciprocal, square root and exponential are the hardware
// instructions (a few ulp) and 1/|x-y| is formed as |x-y| * (1/|x-y|^2); FP64
// keeps exact division, square root and exp.
//
// Multi-indices are stored compactly: degree d = i+j+k, s = j+k,
// index = d(d+1)(d+2)/6 + s(s+1)/2 + k.  All loops have compile-time bounds and
// are fully unrolled so that a[] and b[] live in registers.
#pragma once

namespace tabi {

template <int L>
struct TaylorShape {
    static constexpr int NJ = (L + 1) * (L + 2) * (L + 3) / 6;
};

__host__ __device__ constexpr int tidx(int i, int j, int k)
{
    return (i + j + k) * (i + j + k + 1) * (i + j + k + 2) / 6 + (j + k) * (j + k + 1) / 2 + k;
}

// Inverse of tidx() without loops (sums of comparisons, degrees <= 11), so that
// in a fully unrolled loop over id every use folds to a constant.
__host__ __device__ constexpr int tdeg(int id)
{
    return (id >= 1) + (id >= 4) + (id >= 10) + (id >= 20) + (id >= 35) + (id >= 56) +
           (id >= 84) + (id >= 120) + (id >= 165) + (id >= 220) + (id >= 286);
}
__host__ __device__ constexpr int texp(int id, int axis)
{
    const int d = tdeg(id);
    const int r = id - d * (d + 1) * (d + 2) / 6;             // (j+k)(j+k+1)/2 + k
    const int s = (r >= 1) + (r >= 3) + (r >= 6) + (r >= 10) + (r >= 15) + (r >= 21) +
                  (r >= 28) + (r >= 36) + (r >= 45) + (r >= 55) + (r >= 66);
    const int k = r - s * (s + 1) / 2;
    return axis == 0 ? d - s : axis == 1 ? s - k : k;
}
// compact index of the multi-index id + e_axis
__host__ __device__ constexpr int tshift(int id, int axis)
{
    return tidx(texp(id, 0) + (axis == 0), texp(id, 1) + (axis == 1), texp(id, 2) + (axis == 2));
}

template <typename T>
__device__ __forceinline__ constexpr T cf1(int i) { return T(1) / T(i); }
template <typename T>
__device__ __forceinline__ constexpr T cf2(int i) { return T(1) - T(0.5) * (T(1) / T(i)); }
template <typename T>
__device__ __forceinline__ constexpr T cf3(int i) { return T(1) - T(1) / T(i); }

template <int L, bool KAPPA, typename T>
__device__ __forceinline__ void taylor(T dx, T dy, T dz, T kappa, T (&a)[TaylorShape<L>::NJ])
{
    constexpr int NJ = TaylorShape<L>::NJ;
    T b[NJ];
#define A3(i, j, k) a[tidx(i, j, k)]
#define B3(i, j, k) b[tidx(i, j, k)]
    const T ddx = T(2) * dx, ddy = T(2) * dy, ddz = T(2) * dz;
    const T kx = kappa * dx, ky = kappa * dy, kz = kappa * dz;
    const T dist2 = dx * dx + dy * dy + dz * dz;
    T fac, dist, rdist;
    if constexpr (sizeof(T) == 4) {
        fac = __builtin_amdgcn_rcpf(dist2);
        dist = __builtin_amdgcn_sqrtf(dist2);
        rdist = dist * fac;
    } else {
        fac = T(1) / dist2;
        dist = sqrt(dist2);
        rdist = T(1) / dist;
    }

    if constexpr (KAPPA) {
        if constexpr (sizeof(T) == 4) {
            B3(0, 0, 0) = __expf(-kappa * dist);
            A3(0, 0, 0) = B3(0, 0, 0) * rdist;
        } else {
            B3(0, 0, 0) = exp(-kappa * dist);
            A3(0, 0, 0) = B3(0, 0, 0) / dist;
        }
        B3(1, 0, 0) = kx * A3(0, 0, 0);
        B3(0, 1, 0) = ky * A3(0, 0, 0);
        B3(0, 0, 1) = kz * A3(0, 0, 0);
        const T s0 = A3(0, 0, 0) + kappa * B3(0, 0, 0);
        A3(1, 0, 0) = fac * dx * s0;
        A3(0, 1, 0) = fac * dy * s0;
        A3(0, 0, 1) = fac * dz * s0;
    } else {
        A3(0, 0, 0) = rdist;
        A3(1, 0, 0) = fac * dx * A3(0, 0, 0);
        A3(0, 1, 0) = fac * dy * A3(0, 0, 0);
        A3(0, 0, 1) = fac * dz * A3(0, 0, 0);
    }

#pragma unroll
    for (int i = 2; i <= L; ++i) {
        if constexpr (KAPPA) {
            const T c1 = cf1<T>(i) * kappa;
            B3(i, 0, 0) = c1 * (dx * A3(i - 1, 0, 0) - A3(i - 2, 0, 0));
            B3(0, i, 0) = c1 * (dy * A3(0, i - 1, 0) - A3(0, i - 2, 0));
            B3(0, 0, i) = c1 * (dz * A3(0, 0, i - 1) - A3(0, 0, i - 2));
            A3(i, 0, 0) = fac * (ddx * cf2<T>(i) * A3(i - 1, 0, 0) - cf3<T>(i) * A3(i - 2, 0, 0) +
                                 c1 * (dx * B3(i - 1, 0, 0) - B3(i - 2, 0, 0)));
            A3(0, i, 0) = fac * (ddy * cf2<T>(i) * A3(0, i - 1, 0) - cf3<T>(i) * A3(0, i - 2, 0) +
                                 c1 * (dy * B3(0, i - 1, 0) - B3(0, i - 2, 0)));
            A3(0, 0, i) = fac * (ddz * cf2<T>(i) * A3(0, 0, i - 1) - cf3<T>(i) * A3(0, 0, i - 2) +
                                 c1 * (dz * B3(0, 0, i - 1) - B3(0, 0, i - 2)));
        } else {
            A3(i, 0, 0) = fac * (ddx * cf2<T>(i) * A3(i - 1, 0, 0) - cf3<T>(i) * A3(i - 2, 0, 0));
            A3(0, i, 0) = fac * (ddy * cf2<T>(i) * A3(0, i - 1, 0) - cf3<T>(i) * A3(0, i - 2, 0));
            A3(0, 0, i) = fac * (ddz * cf2<T>(i) * A3(0, 0, i - 1) - cf3<T>(i) * A3(0, 0, i - 2));
        }
    }

    if constexpr (KAPPA) {
        B3(1, 1, 0) = kx * A3(0, 1, 0);
        B3(1, 0, 1) = kx * A3(0, 0, 1);
        B3(0, 1, 1) = ky * A3(0, 0, 1);
        A3(1, 1, 0) = fac * (dx * A3(0, 1, 0) + ddy * A3(1, 0, 0) + kx * B3(0, 1, 0));
        A3(1, 0, 1) = fac * (dx * A3(0, 0, 1) + ddz * A3(1, 0, 0) + kx * B3(0, 0, 1));
        A3(0, 1, 1) = fac * (dy * A3(0, 0, 1) + ddz * A3(0, 1, 0) + ky * B3(0, 0, 1));
    } else {
        A3(1, 1, 0) = fac * (dx * A3(0, 1, 0) + ddy * A3(1, 0, 0));
        A3(1, 0, 1) = fac * (dx * A3(0, 0, 1) + ddz * A3(1, 0, 0));
        A3(0, 1, 1) = fac * (dy * A3(0, 0, 1) + ddz * A3(0, 1, 0));
    }

#pragma unroll
    for (int i = 2; i <= L - 1; ++i) {
        if constexpr (KAPPA) {
            B3(1, 0, i) = kx * A3(0, 0, i);
            B3(0, 1, i) = ky * A3(0, 0, i);
            B3(0, i, 1) = kz * A3(0, i, 0);
            B3(1, i, 0) = kx * A3(0, i, 0);
            B3(i, 1, 0) = ky * A3(i, 0, 0);
            B3(i, 0, 1) = kz * A3(i, 0, 0);
            A3(1, 0, i) = fac * (dx * A3(0, 0, i) + ddz * A3(1, 0, i - 1) - A3(1, 0, i - 2) + kx * B3(0, 0, i));
            A3(0, 1, i) = fac * (dy * A3(0, 0, i) + ddz * A3(0, 1, i - 1) - A3(0, 1, i - 2) + ky * B3(0, 0, i));
            A3(0, i, 1) = fac * (dz * A3(0, i, 0) + ddy * A3(0, i - 1, 1) - A3(0, i - 2, 1) + kz * B3(0, i, 0));
            A3(1, i, 0) = fac * (dx * A3(0, i, 0) + ddy * A3(1, i - 1, 0) - A3(1, i - 2, 0) + kx * B3(0, i, 0));
            A3(i, 1, 0) = fac * (dy * A3(i, 0, 0) + ddx * A3(i - 1, 1, 0) - A3(i - 2, 1, 0) + ky * B3(i, 0, 0));
            A3(i, 0, 1) = fac * (dz * A3(i, 0, 0) + ddx * A3(i - 1, 0, 1) - A3(i - 2, 0, 1) + kz * B3(i, 0, 0));
        } else {
            A3(1, 0, i) = fac * (dx * A3(0, 0, i) + ddz * A3(1, 0, i - 1) - A3(1, 0, i - 2));
            A3(0, 1, i) = fac * (dy * A3(0, 0, i) + ddz * A3(0, 1, i - 1) - A3(0, 1, i - 2));
            A3(0, i, 1) = fac * (dz * A3(0, i, 0) + ddy * A3(0, i - 1, 1) - A3(0, i - 2, 1));
            A3(1, i, 0) = fac * (dx * A3(0, i, 0) + ddy * A3(1, i - 1, 0) - A3(1, i - 2, 0));
            A3(i, 1, 0) = fac * (dy * A3(i, 0, 0) + ddx * A3(i - 1, 1, 0) - A3(i - 2, 1, 0));
            A3(i, 0, 1) = fac * (dz * A3(i, 0, 0) + ddx * A3(i - 1, 0, 1) - A3(i - 2, 0, 1));
        }
    }

#pragma unroll
    for (int i = 2; i <= L - 2; ++i) {
#pragma unroll
        for (int j = 2; j <= L - i; ++j) {
            if constexpr (KAPPA) {
                const T c1 = cf1<T>(i) * kappa;
                B3(i, j, 0) = c1 * (dx * A3(i - 1, j, 0) - A3(i - 2, j, 0));
                B3(i, 0, j) = c1 * (dx * A3(i - 1, 0, j) - A3(i - 2, 0, j));
                B3(0, i, j) = c1 * (dy * A3(0, i - 1, j) - A3(0, i - 2, j));
                A3(i, j, 0) = fac * (ddx * cf2<T>(i) * A3(i - 1, j, 0) + ddy * A3(i, j - 1, 0) -
                                     cf3<T>(i) * A3(i - 2, j, 0) - A3(i, j - 2, 0) +
                                     c1 * (dx * B3(i - 1, j, 0) - B3(i - 2, j, 0)));
                A3(i, 0, j) = fac * (ddx * cf2<T>(i) * A3(i - 1, 0, j) + ddz * A3(i, 0, j - 1) -
                                     cf3<T>(i) * A3(i - 2, 0, j) - A3(i, 0, j - 2) +
                                     c1 * (dx * B3(i - 1, 0, j) - B3(i - 2, 0, j)));
                A3(0, i, j) = fac * (ddy * cf2<T>(i) * A3(0, i - 1, j) + ddz * A3(0, i, j - 1) -
                                     cf3<T>(i) * A3(0, i - 2, j) - A3(0, i, j - 2) +
                                     c1 * (dy * B3(0, i - 1, j) - B3(0, i - 2, j)));
            } else {
                A3(i, j, 0) = fac * (ddx * cf2<T>(i) * A3(i - 1, j, 0) + ddy * A3(i, j - 1, 0) -
                                     cf3<T>(i) * A3(i - 2, j, 0) - A3(i, j - 2, 0));
                A3(i, 0, j) = fac * (ddx * cf2<T>(i) * A3(i - 1, 0, j) + ddz * A3(i, 0, j - 1) -
                                     cf3<T>(i) * A3(i - 2, 0, j) - A3(i, 0, j - 2));
                A3(0, i, j) = fac * (ddy * cf2<T>(i) * A3(0, i - 1, j) + ddz * A3(0, i, j - 1) -
                                     cf3<T>(i) * A3(0, i - 2, j) - A3(0, i, j - 2));
            }
        }
    }

    if constexpr (L >= 3) {
        if constexpr (KAPPA) {
            B3(1, 1, 1) = kx * A3(0, 1, 1);
            A3(1, 1, 1) = fac * (dx * A3(0, 1, 1) + ddy * A3(1, 0, 1) + ddz * A3(1, 1, 0) + kx * B3(0, 1, 1));
        } else {
            A3(1, 1, 1) = fac * (dx * A3(0, 1, 1) + ddy * A3(1, 0, 1) + ddz * A3(1, 1, 0));
        }
    }

#pragma unroll
    for (int i = 2; i <= L - 2; ++i) {
        if constexpr (KAPPA) {
            B3(1, 1, i) = kx * A3(0, 1, i);
            B3(1, i, 1) = kx * A3(0, i, 1);
            B3(i, 1, 1) = ky * A3(i, 0, 1);
            A3(1, 1, i) = fac * (dx * A3(0, 1, i) + ddy * A3(1, 0, i) + ddz * A3(1, 1, i - 1) -
                                 A3(1, 1, i - 2) + kx * B3(0, 1, i));
            A3(1, i, 1) = fac * (dx * A3(0, i, 1) + ddy * A3(1, i - 1, 1) + ddz * A3(1, i, 0) -
                                 A3(1, i - 2, 1) + kx * B3(0, i, 1));
            A3(i, 1, 1) = fac * (dy * A3(i, 0, 1) + ddx * A3(i - 1, 1, 1) + ddz * A3(i, 1, 0) -
                                 A3(i - 2, 1, 1) + ky * B3(i, 0, 1));
        } else {
            A3(1, 1, i) = fac * (dx * A3(0, 1, i) + ddy * A3(1, 0, i) + ddz * A3(1, 1, i - 1) - A3(1, 1, i - 2));
            A3(1, i, 1) = fac * (dx * A3(0, i, 1) + ddy * A3(1, i - 1, 1) + ddz * A3(1, i, 0) - A3(1, i - 2, 1));
            A3(i, 1, 1) = fac * (dy * A3(i, 0, 1) + ddx * A3(i - 1, 1, 1) + ddz * A3(i, 1, 0) - A3(i - 2, 1, 1));
        }
    }

#pragma unroll
    for (int i = 2; i <= L - 3; ++i) {
#pragma unroll
        for (int j = 2; j <= L - 1 - i; ++j) {
            if constexpr (KAPPA) {
                B3(1, i, j) = kx * A3(0, i, j);
                B3(i, 1, j) = ky * A3(i, 0, j);
                B3(i, j, 1) = kz * A3(i, j, 0);
                A3(1, i, j) = fac * (dx * A3(0, i, j) + ddy * A3(1, i - 1, j) + ddz * A3(1, i, j - 1) -
                                     A3(1, i - 2, j) - A3(1, i, j - 2) + kx * B3(0, i, j));
                A3(i, 1, j) = fac * (dy * A3(i, 0, j) + ddx * A3(i - 1, 1, j) + ddz * A3(i, 1, j - 1) -
                                     A3(i - 2, 1, j) - A3(i, 1, j - 2) + ky * B3(i, 0, j));
                A3(i, j, 1) = fac * (dz * A3(i, j, 0) + ddx * A3(i - 1, j, 1) + ddy * A3(i, j - 1, 1) -
                                     A3(i - 2, j, 1) - A3(i, j - 2, 1) + kz * B3(i, j, 0));
            } else {
                A3(1, i, j) = fac * (dx * A3(0, i, j) + ddy * A3(1, i - 1, j) + ddz * A3(1, i, j - 1) -
                                     A3(1, i - 2, j) - A3(1, i, j - 2));
                A3(i, 1, j) = fac * (dy * A3(i, 0, j) + ddx * A3(i - 1, 1, j) + ddz * A3(i, 1, j - 1) -
                                     A3(i - 2, 1, j) - A3(i, 1, j - 2));
                A3(i, j, 1) = fac * (dz * A3(i, j, 0) + ddx * A3(i - 1, j, 1) + ddy * A3(i, j - 1, 1) -
                                     A3(i - 2, j, 1) - A3(i, j - 2, 1));
            }
        }
    }

#pragma unroll
    for (int k = 2; k <= L - 4; ++k) {
#pragma unroll
        for (int j = 2; j <= L - 2 - k; ++j) {
#pragma unroll
            for (int i = 2; i <= L - k - j; ++i) {
                if constexpr (KAPPA) {
                    const T c1 = cf1<T>(i) * kappa;
                    B3(i, j, k) = c1 * (dx * A3(i - 1, j, k) - A3(i - 2, j, k));
                    A3(i, j, k) = fac * (ddx * cf2<T>(i) * A3(i - 1, j, k) + ddy * A3(i, j - 1, k) +
                                         ddz * A3(i, j, k - 1) - cf3<T>(i) * A3(i - 2, j, k) -
                                         A3(i, j - 2, k) - A3(i, j, k - 2) +
                                         c1 * (dx * B3(i - 1, j, k) - B3(i - 2, j, k)));
                } else {
                    A3(i, j, k) = fac * (ddx * cf2<T>(i) * A3(i - 1, j, k) + ddy * A3(i, j - 1, k) +
                                         ddz * A3(i, j, k - 1) - cf3<T>(i) * A3(i - 2, j, k) -
                                         A3(i, j - 2, k) - A3(i, j, k - 2));
                }
            }
        }
    }
#undef A3
#undef B3
    (void)b;
}

}  // namespace tabi
