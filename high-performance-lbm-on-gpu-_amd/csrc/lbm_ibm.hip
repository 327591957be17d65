// Immersed-boundary direct forcing (multi-direct iterations, Roma kernel), fused into one launch.
//
// The reference evaluates this as dozens of array operations per step
// (src/lbm/ibm.py:189-214, ImmersedForcing.force_field). Here one workgroup
// runs the whole pipeline with barriers between phases:
//
//   P0  moments of the streamed populations at the unique stencil nodes
//       (affine rows compiled for the current layout: drho, rho u)
//   P1  rho = 1 + drho,  u* = (rho u + F0/2) / rho,  force = 0
//   P2  rho_l = sum_s w_ls rho(x_ls)                         (per marker)
//   per iteration:
//   P3  u = u* + force / (2 rho)                             (per node)
//   P4  U_l = sum_s w_ls u(x_ls),  F_l = 2 rho_l (0 - U_l) dA_l
//   P5  total += sum_l F_l;  force(x) += sum over entries touching x of w_ls F_l
//   end: field[:, nodes] = force, body_force = -total
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

#include "lbm_common.h"

namespace lbm {

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
// axis is an outer loop of the reduction.
template <class T, class Term>
__device__ __forceinline__ T sequential_sum(int n, Term term)
{
    T s = term(0);
    for (int p = 1; p < n; ++p) s = s + term(p);
    return s;
}

constexpr int IBM_THREADS = 1024;

template <class T>
struct IbmScratch {
    T* moments;   // (d + 1) * nn
    T* rho;       // nn
    T* ustar;     // d * nn
    T* force;     // d * nn
    T* u;         // d * nn
    T* rho_m;     // nm
    T* f_m;       // d * nm
    T* total;     // d
};

template <class T>
__host__ __device__ inline IbmScratch<T> carve(T* base, int d, int nn, int nm)
{
    IbmScratch<T> s;
    s.moments = base;
    s.rho = s.moments + (size_t)(d + 1) * nn;
    s.ustar = s.rho + nn;
    s.force = s.ustar + (size_t)d * nn;
    s.u = s.force + (size_t)d * nn;
    s.rho_m = s.u + (size_t)d * nn;
    s.f_m = s.rho_m + nm;
    s.total = s.f_m + (size_t)d * nm;
    return s;
}

inline size_t ibm_scratch_elems(int d, int nn, int nm) { return (size_t)(d + 1) * nn + nn + 3 * (size_t)d * nn + nm + (size_t)d * nm + d; }

template <class T, int D>
__global__ __launch_bounds__(IBM_THREADS) void k_ibm(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                                     const T* __restrict__ coeff, const T* __restrict__ constant,
                                                     int m, const int64_t* __restrict__ nodes, int nn,
                                                     const int* __restrict__ inverse, const T* __restrict__ weights,
                                                     const T* __restrict__ areas, int nm,
                                                     const int* __restrict__ spread_start,
                                                     const int* __restrict__ spread_entry, int iterations, T hfx, T hfy,
                                                     T hfz, uint32_t n_total, T* __restrict__ field,
                                                     T* __restrict__ body_force, T* __restrict__ scratch)
{
    constexpr int S = D == 2 ? 9 : 27;
    const IbmScratch<T> s = carve(scratch, D, nn, nm);
    const int tid = threadIdx.x;
    const int rows = (D + 1) * nn;
    const T half_force[3] = {hfx, hfy, hfz};

    // P0: affine moment rows, pairwise over the m padded terms.
    for (int r = tid; r < rows; r += IBM_THREADS) {
        const T sum = pairwise_block<T>(m, [&](int p) {
            const int64_t e = (int64_t)p * rows + r;
            const T a = coeff[e];
            return a == T(0) ? T(0) : a * F[slot[e]];
        });
        s.moments[r] = constant[r] + sum;
    }
    __syncthreads();
    // P1
    for (int n = tid; n < nn; n += IBM_THREADS) {
        const T rho = T(1) + s.moments[n];
        s.rho[n] = rho;
        for (int a = 0; a < D; ++a) {
            s.ustar[(size_t)a * nn + n] = (s.moments[(size_t)(1 + a) * nn + n] + half_force[a]) / rho;
            s.force[(size_t)a * nn + n] = T(0);
        }
    }
    if (tid < D) s.total[tid] = T(0);
    __syncthreads();
    // P2
    for (int l = tid; l < nm; l += IBM_THREADS) {
        const int* inv = inverse + (size_t)l * S;
        const T* w = weights + (size_t)l * S;
        s.rho_m[l] = pairwise_block<T>(S, [&](int p) { return w[p] * s.rho[inv[p]]; });
    }
    __syncthreads();
    for (int it = 0; it < iterations; ++it) {
        // P3
        for (int n = tid; n < nn; n += IBM_THREADS) {
            const T two_rho = T(2) * s.rho[n];
            for (int a = 0; a < D; ++a)
                s.u[(size_t)a * nn + n] = s.ustar[(size_t)a * nn + n] + s.force[(size_t)a * nn + n] / two_rho;
        }
        __syncthreads();
        // P4
        for (int l = tid; l < nm; l += IBM_THREADS) {
            const int* inv = inverse + (size_t)l * S;
            const T* w = weights + (size_t)l * S;
            const T two_rho_m = T(2) * s.rho_m[l];
            for (int a = 0; a < D; ++a) {
                const T* ua = s.u + (size_t)a * nn;
                const T um = sequential_sum<T>(S, [&](int p) { return w[p] * ua[inv[p]]; });
                s.f_m[(size_t)a * nm + l] = (two_rho_m * (T(0) - um)) * areas[l];
            }
        }
        __syncthreads();
        // P5: marker totals (one thread per component) and the deterministic spread.
        if (tid < D) {
            const T* fa = s.f_m + (size_t)tid * nm;
            s.total[tid] = s.total[tid] + sequential_sum<T>(nm, [&](int l) { return fa[l]; });
        }
        for (int n = tid; n < nn; n += IBM_THREADS) {
            const int begin = spread_start[n], end = spread_start[n + 1];
            for (int a = 0; a < D; ++a) {
                const T* fa = s.f_m + (size_t)a * nm;
                double acc = 0.0;
                for (int e = begin; e < end; ++e) {
                    const int k = spread_entry[e];
                    const int l = k / S;
                    acc += (double)(weights[k] * fa[l]);
                }
                T& slot_force = s.force[(size_t)a * nn + n];
                slot_force = T((double)slot_force + acc);
            }
        }
        __syncthreads();
    }
    for (int n = tid; n < nn; n += IBM_THREADS)
        for (int a = 0; a < D; ++a) field[(size_t)a * n_total + nodes[n]] = s.force[(size_t)a * nn + n];
    if (tid < D) body_force[tid] = -s.total[tid];
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LBM_API int lbm_ibm_scratch_size(int lattice, int dtype, int n_nodes, int n_markers, int64_t* bytes)
{
    return guarded("lbm_ibm_scratch_size", [&] {
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            *bytes = (int64_t)(ibm_scratch_elems(L::D, n_nodes, n_markers) * sizeof(T));
        });
    });
}

LBM_API int lbm_ibm_force(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                          const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                          const void* weights, const void* areas, int n_markers, const int* spread_start,
                          const int* spread_entry, int iterations, double fx, double fy, double fz, int64_t n_total,
                          void* field, void* body_force, void* scratch, void* stream)
{
    return guarded("lbm_ibm_force", [&] {
        require(n_nodes > 0 && n_markers > 0, "the immersed body needs markers and stencil nodes");
        require(iterations >= 1, "iterations must be >= 1");
        require(m >= 1 && m <= 128, "moment rows need 1 <= m <= 128 terms");
        require(n_total > 0 && n_total < (1ll << 31), "node count out of range");
        require(f && slot && coeff && constant && nodes && inverse && weights && areas && spread_start &&
                    spread_entry && field && body_force && scratch,
                "null pointer");
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const T hf[3] = {T(0.5 * fx), T(0.5 * fy), T(0.5 * fz)};
            k_ibm<T, L::D><<<1, IBM_THREADS, 0, as_stream(stream)>>>(
                static_cast<const T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), m,
                nodes, n_nodes, inverse, static_cast<const T*>(weights), static_cast<const T*>(areas), n_markers,
                spread_start, spread_entry, iterations, hf[0], hf[1], hf[2], (uint32_t)n_total, static_cast<T*>(field),
                static_cast<T*>(body_force), static_cast<T*>(scratch));
            check_launch("IBM kernel");
        });
    });
}

}  // extern "C"
