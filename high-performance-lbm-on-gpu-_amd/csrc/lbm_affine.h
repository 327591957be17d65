// Evaluation of one row of a compiled affine set: constant[r] + sum_t coeff[t][r] F[slot[t][r]].
//
// Term arrays are term-major with row stride `stride` (slot[t * stride + r]). A row's sum is
// NumPy's pairwise summation over its m padded terms (8 interleaved partial sums when
// m >= 8, a plain running sum from 0 otherwise; numpy/_core/src/umath/loops_utils.h.src,
// pairwise_sum for n <= PW_BLOCKSIZE), the order in which the reference evaluates
// (const + (coeff * f[slots]).sum(axis=1)) (src/lbm/link_rules.py:269-272), so a row
// gives the same bits as the NumPy oracle.
//
// Terms are fetched eight at a time: the coefficient and slot loads of a chunk are issued
// together, then the eight population gathers, so a row of m terms costs about
// 2 * ceil(m / 8) memory latencies instead of 2 m. A zero coefficient contributes an exact
// zero whatever the population holds; its gather reads element 0 instead of the padded slot.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_hold.h"

namespace lbm {

template <class T>
__device__ __forceinline__ void affine_terms(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                             const T* __restrict__ coeff, int64_t stride, int64_t r, int p0,
                                             int count, T (&term)[8])
{
    T a[8];
    int64_t s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (k < count) {
            const int64_t e = (int64_t)(p0 + k) * stride + r;
            a[k] = coeff[e];
            s[k] = slot[e];
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (k < count) hold_value(s[k]);  // slot loads stay ahead of the gather branches
    // The gathers go out together from their selected addresses (hold_value: otherwise
    // each would sit behind a branch on its coefficient with a wait of its own). A term
    // whose coefficient is zero on every lane of the wavefront (row padding) is not
    // fetched at all: the branch around it is uniform and its term is zero either way.
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (k < count) v[k] = __any(a[k] != T(0)) ? F[a[k] == T(0) ? 0 : s[k]] : T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (k < count) hold_value(v[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (k < count) term[k] = a[k] == T(0) ? T(0) : a[k] * v[k];
}

// Pairwise sum of the m terms of row r (no constant).
template <class T>
__device__ __forceinline__ T affine_sum(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                        const T* __restrict__ coeff, int64_t stride, int64_t r, int m)
{
    T t[8];
    if (m < 8) {
        affine_terms(F, slot, coeff, stride, r, 0, m, t);
        T s = T(0);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < m) s = s + t[k];
        return s;
    }
    T acc[8];
    affine_terms(F, slot, coeff, stride, r, 0, 8, acc);
    const int lim = m - (m % 8);
    int p = 8;
    for (; p < lim; p += 8) {
        affine_terms(F, slot, coeff, stride, r, p, 8, t);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = acc[k] + t[k];
    }
    T s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    if (p < m) {
        const int rest = m - p;
        affine_terms(F, slot, coeff, stride, r, p, rest, t);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < rest) s = s + t[k];
    }
    return s;
}

template <class T>
__device__ __forceinline__ T affine_row(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                        const T* __restrict__ coeff, const T* __restrict__ constant,
                                        int64_t stride, int m, int64_t r)
{
    const T c = constant[r];
    return c + affine_sum(F, slot, coeff, stride, r, m);
}

// affine_row for rows of at most M terms (M a multiple of 8): all the row's coefficients,
// slots and its constant are fetched at once, then all its populations, so a row costs two
// memory latencies whatever m is. The sum is affine_row's: a running sum from 0 for m < 8,
// else 8 interleaved partial sums over the whole 8-term chunks, their fixed tree, then the
// remaining terms one by one; the same bits.
template <class T, int M>
__device__ __forceinline__ T affine_row_wide(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                             const T* __restrict__ coeff, const T* __restrict__ constant,
                                             int64_t stride, int m, int64_t r)
{
    static_assert(M % 8 == 0 && M >= 8, "M is a multiple of 8");
    T a[M];
    int64_t s[M];
    const T c = constant[r];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if (k < m) {
            const int64_t e = (int64_t)k * stride + r;
            a[k] = coeff[e];
            s[k] = slot[e];
        }
    }
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < m) hold_value(s[k]);
    T v[M];
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < m) v[k] = F[a[k] == T(0) ? 0 : s[k]];
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < m) hold_value(v[k]);
    T t[M];
#pragma unroll
    for (int k = 0; k < M; ++k) t[k] = (k < m && a[k] != T(0)) ? a[k] * v[k] : T(0);
    T sum = T(0);
    if (m < 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < m) sum = sum + t[k];
        return c + sum;
    }
    const int lim = m - (m % 8);
    T acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = t[k];
#pragma unroll
    for (int p = 8; p < M; p += 8) {
        if (p < lim) {
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] = acc[k] + t[p + k];
        }
    }
    sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
    for (int p = 8; p < M; ++p)
        if (p >= lim && p < m) sum = sum + t[p];
    return c + sum;
}

// One row's coefficients, slots and constant fetched ahead (rows of at most M terms, 8 <= M
// <= 16), so a kernel can issue those loads before other independent work and gather
// from F afterwards; affine_finish then sums in affine_row's order (8 interleaved partial
// sums over whole 8-term chunks, their fixed tree, then the rest one by one), giving the
// same bits.
template <class T, int M>
struct AffineAhead {
    static_assert(M >= 8 && M <= 16, "rows of 8 to 16 terms");
    T a[M];
    int64_t s[M];
    T c;
};

template <class T, int M>
__device__ __forceinline__ void affine_fetch(const int64_t* __restrict__ slot, const T* __restrict__ coeff,
                                             const T* __restrict__ constant, int64_t stride, int m, int64_t r,
                                             AffineAhead<T, M>& p)
{
    p.c = constant[r];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if (k < m) {
            const int64_t e = (int64_t)k * stride + r;
            p.a[k] = coeff[e];
            p.s[k] = slot[e];
        }
    }
}

template <class T, int M>
__device__ __forceinline__ T affine_finish(const T* __restrict__ F, const AffineAhead<T, M>& p, int m)
{
    T v[M], t[M];
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < m) v[k] = F[p.a[k] == T(0) ? 0 : p.s[k]];
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < m) hold_value(v[k]);
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < m) t[k] = p.a[k] == T(0) ? T(0) : p.a[k] * v[k];
    T s = T(0);
    if (m < 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < m) s = s + t[k];
        return p.c + s;
    }
    T acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = t[k];
    const int rest = (M == 16 && m == 16) ? 16 : 8;  // a second whole chunk only when m == 16
    if constexpr (M == 16) {
        if (m == 16) {
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] = acc[k] + t[8 + k];
        }
    }
    s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
    for (int k = 8; k < M; ++k)
        if (k >= rest && k < m) s = s + t[k];
    return p.c + s;
}

// The link-rule pre-pass by one workgroup, in two halves so a kernel can do other reads of
// the state between them: rules_values evaluates rows r = tid + k * nthreads (k < PER, so
// at most PER * nthreads rules) into registers and fetches their targets; after a barrier,
// rules_store writes F[target[r]], so every rule reads pre-rule populations.
template <class T, int PER>
__device__ __forceinline__ void rules_values(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                             const T* __restrict__ coeff, const T* __restrict__ constant,
                                             const int64_t* __restrict__ target, int64_t n, int m, int tid,
                                             int nthreads, T (&v)[PER], int64_t (&t)[PER])
{
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int64_t r = (int64_t)k * nthreads + tid;
        if (r < n) t[k] = target[r];  // fetched with the terms, used after the barrier
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int64_t r = (int64_t)k * nthreads + tid;
        v[k] = r < n ? affine_row(F, slot, coeff, constant, n, m, r) : T(0);
    }
}

template <class T, int PER>
__device__ __forceinline__ void rules_store(T* __restrict__ F, int64_t n, int tid, int nthreads, const T (&v)[PER],
                                            const int64_t (&t)[PER])
{
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int64_t r = (int64_t)k * nthreads + tid;
        if (r < n) F[t[k]] = v[k];
    }
}

// The whole pre-pass by one workgroup. Call from all lanes of the group.
template <class T, int PER>
__device__ __forceinline__ void rules_in_one_group(T* __restrict__ F, const int64_t* __restrict__ slot,
                                                   const T* __restrict__ coeff, const T* __restrict__ constant,
                                                   const int64_t* __restrict__ target, int64_t n, int m, int tid,
                                                   int nthreads)
{
    T v[PER];
    int64_t t[PER];
    rules_values<T, PER>(F, slot, coeff, constant, target, n, m, tid, nthreads, v, t);
    __syncthreads();
    rules_store<T, PER>(F, n, tid, nthreads, v, t);
}

}  // namespace lbm
