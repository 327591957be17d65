// Multi-direct immersed-boundary forcing (Roma kernel), the part shared by the forward IBM
// kernel and its adjoint: the layout of the intermediates and the iteration phases.
//
// The reference evaluates the forcing as array operations (src/lbm/ibm.py:189-214,
// ImmersedForcing.force_field). Here one workgroup runs it with barriers between phases:
//
//   P0  moments of the streamed populations at the unique stencil nodes (affine rows)
//   P1  rho = 1 + drho,  u* = (rho u + F0/2) / rho,  force = 0
//   P2  rho_l = sum_s w_ls rho(x_ls)                         (per marker)
//   per iteration:
//   P3  u = u* + force / (2 rho)                             (per node)
//   P4  U_l = sum_s w_ls u(x_ls),  F_l = 2 rho_l (0 - U_l) dA_l
//   P5  total += sum_l F_l;  force(x) += sum over entries touching x of w_ls F_l
//
// Summation orders are the ones NumPy uses for the reference's expressions:
// pairwise over the padded terms of a moment row and over the stencil for
// rho_l (contiguous row sums), but sequential over the stencil for U_l and over
// the markers for the total, because u[:, inverse] is laid out with the
// component axis innermost and NumPy then reduces those axes as outer loops;
// the spread accumulates in float64 in flat (marker, stencil) order from 0 and
// is added to the float64-widened force, as numpy.bincount does. Each node
// walks its own precomputed entry list, so the result is deterministic
// without atomics.
//
// Layout of the intermediates (nn stencil nodes, nm markers, `slots` iteration slots, d
// components), which is also the per-step record the adjoint reads back:
//   moments (d+1) nn | rho nn | ustar d nn | rho_m nm |
//   force (slots+1) d nn   (force before iteration it, and after the last)
//   u slots d nn | U slots d nm | Fm slots d nm
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_hold.h"

#include <cstddef>
#include <cstdint>

namespace lbm {

// NumPy's pairwise summation of term(0..m-1) for m <= 128.
template <class T, class Term>
__device__ __forceinline__ T pairwise_block(int m, Term term)
{
    if (m < 8) {
        T s = T(0);
        for (int p = 0; p < m; ++p) s = s + term(p);
        return s;
    }
    T r0 = term(0), r1 = term(1), r2 = term(2), r3 = term(3);
    T r4 = term(4), r5 = term(5), r6 = term(6), r7 = term(7);
    const int lim = m - (m % 8);
    int p = 8;
    for (; p < lim; p += 8) {
        r0 = r0 + term(p + 0);
        r1 = r1 + term(p + 1);
        r2 = r2 + term(p + 2);
        r3 = r3 + term(p + 3);
        r4 = r4 + term(p + 4);
        r5 = r5 + term(p + 5);
        r6 = r6 + term(p + 6);
        r7 = r7 + term(p + 7);
    }
    T s = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; p < m; ++p) s = s + term(p);
    return s;
}

// Sequential sum term(0) + term(1) + ..., the order NumPy uses when the reduced
// axis is an outer loop of the reduction. Terms are fetched eight at a time before
// they are added in order, so a long sum costs one memory latency per eight terms.
template <class T, class Term>
__device__ __forceinline__ T sequential_sum(int n, Term term)
{
    T s = term(0);
    int p = 1;
    for (; p + 8 <= n; p += 8) {
        T v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = term(p + k);
#pragma unroll
        for (int k = 0; k < 8; ++k) s = s + v[k];
    }
    for (; p < n; ++p) s = s + term(p);
    return s;
}

// Copies of several tables at once. Each lane issues the loads of every table for its next
// indices before it stores any, so staging them costs about one memory latency instead of
// one per table. The WordCopy list moves 4-byte words; the extra `vec` copy moves 16-byte
// units when both of its ends are 16-byte aligned (the last words that do not fill a unit,
// or the whole copy when misaligned, go word by word).
struct WordCopy {
    uint32_t* dst;
    const uint32_t* src;
    size_t words;
};

template <class T, class U>
__device__ __forceinline__ WordCopy word_copy(T* dst, const U* src, size_t count)
{
    static_assert(sizeof(T) == sizeof(U) && sizeof(T) % 4 == 0, "word copies of equal 4-byte multiples");
    return WordCopy{reinterpret_cast<uint32_t*>(dst), reinterpret_cast<const uint32_t*>(src),
                    count * (sizeof(T) / 4)};
}

// Source of the copies of empty arrays: the loads below are issued unconditionally. Never
// written; not const, so it sits in the global address space like the arrays it stands in for
// (a constant-space stand-in made every source pointer generic).
static __device__ __align__(16) uint32_t copy_zero[4] = {0, 0, 0, 0};

// Copies of N word tables and NV further arrays moved in 16-byte units (KV units of each per
// lane and round; words that do not fill a unit, or a whole array whose ends are not both
// 16-byte aligned, go word by word). Every lane loads from an index valid either way (the
// last element past an array's end, a zero word for an empty array) and only the stores are
// guarded, and the first round is not inside the loop, so all loads of a round go out
// together, also with earlier loads of the caller still in flight: guarded loads compiled to
// one branch and one full wait each, and a load target written before its load (the value of
// an empty array) or a loop header made the compiler wait for every earlier load.
template <int N, int NV, int KV = 2>
__device__ __forceinline__ void copy_staged(const WordCopy (&c)[N], const WordCopy (&vec)[NV], int tid, int nthreads)
{
    using V4 = uint32_t __attribute__((ext_vector_type(4)));
    constexpr int KW = 2;
    const size_t step = (size_t)nthreads;
    size_t units[NV];
    WordCopy rest[NV];
    size_t rounds = 0, most = 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const bool aligned =
            ((reinterpret_cast<uintptr_t>(vec[j].dst) | reinterpret_cast<uintptr_t>(vec[j].src)) & 15) == 0;
        units[j] = aligned ? vec[j].words / 4 : 0;
        rest[j] = WordCopy{vec[j].dst + units[j] * 4, vec[j].src + units[j] * 4, vec[j].words - units[j] * 4};
        const size_t r = (units[j] + KV * step - 1) / (KV * step);
        rounds = r > rounds ? r : rounds;
        most = rest[j].words > most ? rest[j].words : most;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) most = c[j].words > most ? c[j].words : most;
    const size_t word_rounds = (most + KW * step - 1) / (KW * step);
    if (word_rounds > rounds) rounds = word_rounds;
    const V4* vsrc[NV];
    size_t vlast[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        vsrc[j] = units[j] ? reinterpret_cast<const V4*>(vec[j].src) : reinterpret_cast<const V4*>(copy_zero);
        vlast[j] = units[j] ? units[j] - 1 : 0;
    }
    const uint32_t* wsrc[N + NV];
    size_t wlast[N + NV];
#pragma unroll
    for (int j = 0; j < N + NV; ++j) {
        const WordCopy& cj = j < N ? c[j] : rest[j - N];
        wsrc[j] = cj.words ? cj.src : copy_zero;
        wlast[j] = cj.words ? cj.words - 1 : 0;
    }
    auto round = [&](size_t r) {
        const size_t iw = tid + r * KW * step, iv = tid + r * KV * step;
        V4 v[NV][KV];
        uint32_t w[N + NV][KW];
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
            for (int k = 0; k < KV; ++k) {
                const size_t i = iv + k * step;
                v[j][k] = vsrc[j][i < vlast[j] ? i : vlast[j]];
            }
#pragma unroll
        for (int j = 0; j < N + NV; ++j)
#pragma unroll
            for (int k = 0; k < KW; ++k) {
                const size_t i = iw + k * step;
                w[j][k] = wsrc[j][i < wlast[j] ? i : wlast[j]];
            }
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            V4* dst = reinterpret_cast<V4*>(vec[j].dst);
#pragma unroll
            for (int k = 0; k < KV; ++k)
                if (iv + k * step < units[j]) dst[iv + k * step] = v[j][k];
        }
#pragma unroll
        for (int j = 0; j < N + NV; ++j) {
            const WordCopy& cj = j < N ? c[j] : rest[j - N];
#pragma unroll
            for (int k = 0; k < KW; ++k)
                if (iw + k * step < cj.words) cj.dst[iw + k * step] = w[j][k];
        }
    };
    if (rounds > 0) round(0);
    for (size_t r = 1; r < rounds; ++r) round(r);
}

// Copy of `words` 4-byte words from LDS to global memory, in 16-byte units when both ends
// are 16-byte aligned (the forward kernel's record).
__device__ __forceinline__ void copy_from_lds(uint32_t* dst, const uint32_t* src, size_t words, int tid, int nthreads)
{
    using V4 = uint32_t __attribute__((ext_vector_type(4)));
    const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0;
    const size_t units = aligned ? words / 4 : 0;
    for (size_t i = tid; i < units; i += nthreads) reinterpret_cast<V4*>(dst)[i] = reinterpret_cast<const V4*>(src)[i];
    for (size_t i = units * 4 + tid; i < words; i += nthreads) dst[i] = src[i];
}

// N word tables and one array in 16-byte units (copy_staged with one unit array).
template <int N>
__device__ __forceinline__ void copy_tables(const WordCopy (&c)[N], WordCopy vec, int tid, int nthreads)
{
    const WordCopy vecs[1] = {vec};
    copy_staged<N, 1>(c, vecs, tid, nthreads);
}

template <class T>
struct IbmForward {
    T *moments, *rho, *ustar, *rho_m;
    T *force, *u, *U, *Fm;
};

// Parts are carved from `base` in order and stored into *s when s is given. The stores sit
// behind a test of s, not behind a reference that selects between *s and a dummy: the
// select would keep the caller's pointer struct in private memory, and every access through
// it would become a generic (flat) access instead of an LDS one.
template <class T>
__host__ __device__ __forceinline__ size_t ibm_forward_layout(T* base, int d, int nn, int nm, int slots, IbmForward<T>* s)
{
    size_t o = 0;
    auto take = [&](size_t count) {
        T* p = base + o;
        o += count;
        return p;
    };
    T* moments = take((size_t)(d + 1) * nn);
    T* rho = take(nn);
    T* ustar = take((size_t)d * nn);
    T* rho_m = take(nm);
    T* force = take((size_t)(slots + 1) * d * nn);
    T* u = take((size_t)slots * d * nn);
    T* U = take((size_t)slots * d * nm);
    T* Fm = take((size_t)slots * d * nm);
    if (s) *s = IbmForward<T>{moments, rho, ustar, rho_m, force, u, U, Fm};
    return o;
}

__host__ __device__ inline size_t ibm_record_elems(int d, int nn, int nm, int slots)
{
    return ibm_forward_layout<float>(nullptr, d, nn, nm, slots, nullptr);
}

// Stencil tables of the body: inverse[nm][S] (stencil entry -> node), weights[nm][S],
// areas[nm], and per node the increasing list spread_entry[spread_start[n] ..
// spread_start[n + 1]) of flat (marker, stencil) entries that touch it.
template <class T>
struct IbmTables {
    const int* inverse;
    const T* weights;
    const T* areas;
    const int* spread_start;
    const int* spread_entry;
};

// P1..P5 after P0 has filled s.moments (and a barrier). With every_slot, iteration it
// uses slot it of u, U, Fm and reads force slot it, writing slot it + 1; otherwise all
// iterations share slot 0 and the force update is in place. `total` (d values, may be
// null) accumulates the marker forces; its lanes are the last d threads of the group.
// Ends with a barrier; the final force is slot (every_slot ? iterations : 0).
template <class T, int D>
__device__ __forceinline__ void ibm_iterate(const IbmForward<T> s, const IbmTables<T> tb, int nn, int nm,
                                            int iterations, bool every_slot, const T (&half_force)[3],
                                            T* total, int tid, int nthreads)
{
    constexpr int S = D == 2 ? 9 : 27;
    const size_t node_stride = every_slot ? (size_t)D * nn : 0;
    const size_t marker_stride = every_slot ? (size_t)D * nm : 0;
    // P1
    for (int n = tid; n < nn; n += nthreads) {
        const T rho = T(1) + s.moments[n];
        s.rho[n] = rho;
        for (int a = 0; a < D; ++a) {
            s.ustar[(size_t)a * nn + n] = (s.moments[(size_t)(1 + a) * nn + n] + half_force[a]) / rho;
            s.force[(size_t)a * nn + n] = T(0);
        }
    }
    if (total && tid < D) total[tid] = T(0);
    __syncthreads();
    // P2
    for (int l = tid; l < nm; l += nthreads) {
        const int* inv = tb.inverse + (size_t)l * S;
        const T* w = tb.weights + (size_t)l * S;
        s.rho_m[l] = pairwise_block<T>(S, [&](int p) { return w[p] * s.rho[inv[p]]; });
    }
    __syncthreads();
    const int total_lane = tid - (nthreads - D);
    for (int it = 0; it < iterations; ++it) {
        const T* fo = s.force + it * node_stride;
        T* fn = s.force + (it + 1) * node_stride;
        T* u = s.u + it * node_stride;
        T* U = s.U + it * marker_stride;
        T* Fm = s.Fm + it * marker_stride;
        // P3
        for (int n = tid; n < nn; n += nthreads) {
            const T two_rho = T(2) * s.rho[n];
            for (int a = 0; a < D; ++a)
                u[(size_t)a * nn + n] = s.ustar[(size_t)a * nn + n] + fo[(size_t)a * nn + n] / two_rho;
        }
        __syncthreads();
        // P4
        for (int l = tid; l < nm; l += nthreads) {
            int inv[S];
            T w[S];
#pragma unroll
            for (int p = 0; p < S; ++p) {
                inv[p] = tb.inverse[(size_t)l * S + p];
                w[p] = tb.weights[(size_t)l * S + p];
            }
            const T two_rho_m = T(2) * s.rho_m[l];
            const T area = tb.areas[l];
            for (int a = 0; a < D; ++a) {
                const T* ua = u + (size_t)a * nn;
                const T um = sequential_sum<T>(S, [&](int p) { return w[p] * ua[inv[p]]; });
                U[(size_t)a * nm + l] = um;
                Fm[(size_t)a * nm + l] = (two_rho_m * (T(0) - um)) * area;
            }
        }
        __syncthreads();
        // P5: marker totals and the deterministic spread (all components per entry walk).
        if (total && total_lane >= 0) {
            const T* fa = Fm + (size_t)total_lane * nm;
            total[total_lane] = total[total_lane] + sequential_sum<T>(nm, [&](int l) { return fa[l]; });
        }
        for (int n = tid; n < nn; n += nthreads) {
            const int begin = tb.spread_start[n], end = tb.spread_start[n + 1];
            double acc[D];
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] = 0.0;
            for (int e = begin; e < end; ++e) {
                const int k = tb.spread_entry[e];
                const int l = k / S;
                const T w = tb.weights[k];
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += (double)(w * Fm[(size_t)a * nm + l]);
            }
#pragma unroll
            for (int a = 0; a < D; ++a) fn[(size_t)a * nn + n] = T((double)fo[(size_t)a * nn + n] + acc[a]);
        }
        __syncthreads();
    }
}

}  // namespace lbm
