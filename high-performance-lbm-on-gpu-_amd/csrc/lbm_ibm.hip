// Immersed-boundary direct forcing, all multi-direct iterations fused into one launch of
// one workgroup (phases, summation orders and layout in lbm_ibm_core.h). Replaces the
// reference's array-code evaluation (src/lbm/ibm.py:189-214, ImmersedForcing.force_field)
// and its field scatter (src/lbm/backends/cuda_solver.py:286-292).

#include "lbm_affine.h"
#include "lbm_common.h"
#include "lbm_ibm_core.h"

namespace lbm {

// Workgroup bound: 1024 lanes in 2D; 512 in 3D, where the 27-point stencil rows held in
// registers need more than the 128 VGPRs a 1024-lane group leaves each lane.
constexpr int ibm_max_threads(int d) { return d == 2 ? 1024 : 512; }

// Intermediates: the forward layout with `slots` iteration slots (every iteration when
// recording, one reused slot otherwise), then the d running marker totals.
__host__ __device__ inline size_t ibm_state_elems(int d, int nn, int nm, int slots)
{
    return ibm_record_elems(d, nn, nm, slots) + d;
}

// Bytes of LDS the staged variant needs: the intermediates plus the stencil tables.
inline size_t ibm_lds_bytes(int d, int nn, int nm, int slots, size_t real)
{
    const int S = d == 2 ? 9 : 27;
    const size_t reals = ibm_state_elems(d, nn, nm, slots) + (size_t)nm * S + nm;
    const size_t ints = (size_t)nm * S * 2 + nn + 1;
    return reals * real + ints * sizeof(int) + 16;
}

// Workgroup size: enough lanes for one moment row each, in whole wavefronts, from 256 up
// to the bound.
inline int ibm_threads(int d, int nn)
{
    const int rows = (d + 1) * nn;
    return std::max(256, std::min(ibm_max_threads(d), 64 * ((rows + 63) / 64)));
}

// STAGED: intermediates and the stencil tables (inverse, weights, areas, spread lists)
// live in LDS, so every phase after the moment gather runs at LDS latency; otherwise they
// stay in global memory (large bodies). RECORD: every iteration keeps its own u, U, F_l and
// force slot and the forward part of the layout is copied to `record` at the end (written
// in place when not staged), for the adjoint to read back. The arithmetic is the same in
// all four variants.
template <class T, int D, bool STAGED, bool RECORD>
__global__ __launch_bounds__(ibm_max_threads(D)) void k_ibm(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                                     const T* __restrict__ coeff, const T* __restrict__ constant,
                                                     int m, const int64_t* __restrict__ nodes, int nn,
                                                     const int* __restrict__ inverse_g, const T* __restrict__ weights_g,
                                                     const T* __restrict__ areas_g, int nm,
                                                     const int* __restrict__ start_g,
                                                     const int* __restrict__ entry_g, int iterations, T hfx, T hfy,
                                                     T hfz, uint32_t n_total, T* __restrict__ field,
                                                     T* __restrict__ body_force, T* __restrict__ scratch,
                                                     T* __restrict__ record)
{
    constexpr int S = D == 2 ? 9 : 27;
    extern __shared__ __align__(16) unsigned char ibm_lds[];
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int slots = RECORD ? iterations : 1;
    IbmForward<T> s;
    IbmTables<T> tb{inverse_g, weights_g, areas_g, start_g, entry_g};
    T* total;
    T* base;
    const int rows = (D + 1) * nn;
    // With one moment row per lane (2D), the row's coefficients and slots go out before the
    // table copies, so its population gathers wait one latency after the tables, not two.
    constexpr int AHEAD = 9;
    const bool ahead = D == 2 && m <= AHEAD && rows <= nthreads;
    AffineAhead<T, AHEAD> pre;
    if constexpr (D == 2) {
        if (ahead && tid < rows) affine_fetch(slot, coeff, constant, rows, m, (int64_t)tid, pre);
    }
    if constexpr (STAGED) {
        base = reinterpret_cast<T*>(ibm_lds);
        total = base + ibm_forward_layout(base, D, nn, nm, slots, &s);
        T* w = total + D;
        T* a = w + (size_t)nm * S;
        int* inv = reinterpret_cast<int*>(a + nm);
        int* ent = inv + (size_t)nm * S;
        int* st = ent + (size_t)nm * S;
        const WordCopy tables[5] = {word_copy(inv, inverse_g, (size_t)nm * S), word_copy(w, weights_g, (size_t)nm * S),
                                    word_copy(ent, entry_g, (size_t)nm * S), word_copy(a, areas_g, (size_t)nm),
                                    word_copy(st, start_g, (size_t)nn + 1)};
        copy_tables(tables, WordCopy{nullptr, nullptr, 0}, tid, nthreads);
        tb = IbmTables<T>{inv, w, a, st, ent};
    } else {
        base = RECORD ? record : scratch;
        const size_t used = ibm_forward_layout(base, D, nn, nm, slots, &s);
        total = RECORD ? scratch : base + used;
    }
    const T half_force[3] = {hfx, hfy, hfz};
    // P0: affine moment rows, pairwise over the m padded terms.
    if (ahead) {
        if (tid < rows) s.moments[tid] = affine_finish(F, pre, m);
    } else {
        for (int r = tid; r < rows; r += nthreads) s.moments[r] = affine_row(F, slot, coeff, constant, rows, m, r);
    }
    __syncthreads();
    ibm_iterate<T, D>(s, tb, nn, nm, iterations, RECORD, half_force, total, tid, nthreads);
    const T* last = s.force + (RECORD ? (size_t)iterations * D * nn : 0);
    for (int n = tid; n < nn; n += nthreads)
        for (int a = 0; a < D; ++a) field[(size_t)a * n_total + nodes[n]] = last[(size_t)a * nn + n];
    if (tid < D) body_force[tid] = -total[tid];
    if constexpr (STAGED && RECORD) {
        const WordCopy none[1] = {WordCopy{nullptr, nullptr, 0}};
        copy_tables(none, word_copy(record, base, ibm_record_elems(D, nn, nm, iterations)), tid, nthreads);
    }
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
            *bytes = (int64_t)(ibm_state_elems(L::D, n_nodes, n_markers, 1) * sizeof(T));
        });
    });
}

LBM_API int lbm_ibm_record_size(int lattice, int dtype, int n_nodes, int n_markers, int iterations, int64_t* bytes)
{
    return guarded("lbm_ibm_record_size", [&] {
        require(n_nodes > 0 && n_markers > 0 && iterations >= 1, "bad IBM sizes");
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            // rounded up to 16 bytes, so consecutive records keep 16-byte alignment
            const size_t raw = ibm_record_elems(L::D, n_nodes, n_markers, iterations) * sizeof(T);
            *bytes = (int64_t)((raw + 15) / 16 * 16);
        });
    });
}

LBM_API int lbm_ibm_force(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                          const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                          const void* weights, const void* areas, int n_markers, const int* spread_start,
                          const int* spread_entry, int iterations, double fx, double fy, double fz, int64_t n_total,
                          void* field, void* body_force, void* scratch, void* record, void* stream)
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
            const size_t lds = ibm_lds_bytes(L::D, n_nodes, n_markers, record ? iterations : 1, sizeof(T));
            auto go = [&](auto staged, auto recorded) {
                constexpr bool ST = decltype(staged)::value;
                constexpr bool RE = decltype(recorded)::value;
                if (ST) allow_lds(k_ibm<T, L::D, ST, RE>, lds);
                k_ibm<T, L::D, ST, RE><<<1, ibm_threads(L::D, n_nodes), ST ? lds : 0, as_stream(stream)>>>(
                    static_cast<const T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), m,
                    nodes, n_nodes, inverse, static_cast<const T*>(weights), static_cast<const T*>(areas),
                    n_markers, spread_start, spread_entry, iterations, hf[0], hf[1], hf[2], (uint32_t)n_total,
                    static_cast<T*>(field), static_cast<T*>(body_force), static_cast<T*>(scratch),
                    static_cast<T*>(record));
            };
            const bool staged = lds <= lds_limit();
            if (staged && record) go(std::true_type{}, std::true_type{});
            else if (staged) go(std::true_type{}, std::false_type{});
            else if (record) go(std::false_type{}, std::true_type{});
            else go(std::false_type{}, std::false_type{});
            check_launch("IBM kernel");
        });
    });
}

}  // extern "C"
