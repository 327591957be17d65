// Immersed-boundary direct forcing, all multi-direct iterations fused into one launch of
// one workgroup (phases, summation orders and layout in lbm_ibm_core.h). Replaces the
// reference's array-code evaluation (src/lbm/ibm.py:189-214, ImmersedForcing.force_field)
// and its field scatter (src/lbm/backends/cuda_solver.py:286-292).

#include "lbm_affine.h"
#include "lbm_common.h"
#include "lbm_ibm_core.h"
#include "lbm_rule_layer.h"

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
// all four variants. `before` runs ahead of the first loads (a caller's own loads go out
// with them) and `after(last)` once the field is written, `last` being the final force at
// the stencil nodes (component a of node n at last[a nn + n]). With ROWS the caller's
// `rows(moments)` writes the moment rows' values itself (the rows are then neither read
// nor gathered here).
struct IbmNoHook {
    __device__ void operator()() const {}
    template <class U>
    __device__ void operator()(const U*) const
    {
    }
};

template <class T, int D, bool STAGED, bool RECORD, class Before = IbmNoHook, class After = IbmNoHook,
          bool ROWS = false, class Rows = IbmNoHook>
__device__ __forceinline__ void ibm_force_group(const T* __restrict__ F, const int64_t* __restrict__ slot,
                                                const T* __restrict__ coeff, const T* __restrict__ constant, int m,
                                                const int64_t* __restrict__ nodes, int nn,
                                                const int* __restrict__ inverse_g, const T* __restrict__ weights_g,
                                                const T* __restrict__ areas_g, int nm, const int* __restrict__ start_g,
                                                const int* __restrict__ entry_g, int iterations, T hfx, T hfy, T hfz,
                                                uint32_t n_total, T* __restrict__ field, T* __restrict__ body_force,
                                                T* __restrict__ scratch, T* __restrict__ record, int tid, int nthreads,
                                                Before before = Before{}, After after = After{}, Rows rows_hook = Rows{})
{
    constexpr int S = D == 2 ? 9 : 27;
    extern __shared__ __align__(16) unsigned char ibm_lds[];
    const int slots = RECORD ? iterations : 1;
    IbmForward<T> s;
    IbmTables<T> tb{inverse_g, weights_g, areas_g, start_g, entry_g};
    T* total;
    T* base;
    const int rows = (D + 1) * nn;
    before();
    // With one moment row per lane (2D), the row's coefficients and slots go out before the
    // table copies, so its population gathers wait one latency after the tables, not two.
    constexpr int AHEAD = 9;
    const bool ahead = !ROWS && D == 2 && m <= AHEAD && rows <= nthreads;
    AffineAhead<T, AHEAD> pre;
    if constexpr (D == 2 && !ROWS) {
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
    if constexpr (ROWS) {
        rows_hook(s.moments);
    } else if (ahead) {
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
    after(last);
    if constexpr (STAGED && RECORD) {
        const WordCopy out = word_copy(record, base, ibm_record_elems(D, nn, nm, iterations));
        copy_from_lds(out.dst, out.src, out.words, tid, nthreads);
    }
}

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
    ibm_force_group<T, D, STAGED, RECORD>(F, slot, coeff, constant, m, nodes, nn, inverse_g, weights_g, areas_g, nm,
                                          start_g, entry_g, iterations, hfx, hfy, hfz, n_total, field, body_force,
                                          scratch, record, (int)threadIdx.x, (int)blockDim.x);
}

// ------------------------------------------------------------------ forcing inside the step
//
// One forward step of a body-forced lattice in one launch: the IBM forcing of the state
// `src` and the pull step from `src` to `dst` with that forcing (what lbm_ibm_force followed
// by lbm_pull_step_links computes; one-dimensional grid, one workgroup shape):
//   workgroup 0     the forcing (ibm_force_group, staged) and then the pull update of the IBM
//                   stencil nodes, node n by lane n: their populations are loaded with the
//                   forcing's first loads and their force is the one just computed, read
//                   from LDS (the value the field then holds);
//   next `chunks`   the link-rule nodes (rule_chunk, as k_pull_links);
//   the rest        the pull update of every other node (tiles; the stencil nodes carry
//                   KIND_ADJ_REGION in this launch's kind bytes and are skipped).
// Only the stencil nodes are forced (KIND_FORCED), so no other workgroup reads the field the
// forcing writes; none has a solid upstream neighbour or link rules (the host checks), so
// their update is the plain pull with the same operations as pull_update: the same bits.
template <class L, class T, ForceMode FM, bool RECORD>
__global__ __launch_bounds__(1024) void k_ibm_step(const T* __restrict__ src, const int64_t* __restrict__ slot,
                                                   const T* __restrict__ coeff, const T* __restrict__ constant, int m,
                                                   const int64_t* __restrict__ nodes, int nn,
                                                   const int* __restrict__ inverse_g, const T* __restrict__ weights_g,
                                                   const T* __restrict__ areas_g, int nm,
                                                   const int* __restrict__ start_g, const int* __restrict__ entry_g,
                                                   int iterations, T hfx, T hfy, T hfz, T* __restrict__ field,
                                                   T* __restrict__ body_force, T* __restrict__ record,
                                                   T* __restrict__ dst, const uint8_t* __restrict__ kind,
                                                   const T* __restrict__ wall_u, const T* __restrict__ body, Grid g,
                                                   Relaxation<T> r, RuleRows<T> rules, int gx, int gy, int own_rows)
{
    const int b = (int)blockIdx.x;
    const int lanes = (int)(blockDim.x * blockDim.y);
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    if (b == 0) {
        const bool has = tid < nn;
        uint32_t node = 0;
        uint8_t k = 0;
        T f[L::Q], bd[3];
        auto before = [&] {
            if (!has) return;
            node = (uint32_t)nodes[tid];
            const uint32_t rest = node / (uint32_t)g.nx;
            const int x = (int)(node - rest * (uint32_t)g.nx), y = (int)(rest % (uint32_t)g.ny),
                      z = (int)(rest / (uint32_t)g.ny);
            k = kind[node];
            static_for<0, L::Q>([&](auto I) { f[I] = src[(size_t)I * g.ns + neighbour<L, I, -1>(x, y, z, g)]; });
#pragma unroll
            for (int a = 0; a < L::D; ++a) bd[a] = FM == ForceMode::DENSE ? body[(size_t)a * g.n + node] : T(0);
        };
        auto after = [&](const T* last) {
            if (!has) return;
            // node_force with the field value of this node taken from the forcing
            T Fx = r.force[0], Fy = r.force[1], Fz = r.force[2];
            if constexpr (FM == ForceMode::DENSE) {
                Fx = Fx + bd[0];
                Fy = Fy + bd[1];
                if constexpr (L::D == 3) Fz = Fz + bd[2];
            }
            if (k & KIND_FORCED) {
                Fx = Fx + last[tid];
                Fy = Fy + last[(size_t)nn + tid];
                if constexpr (L::D == 3) Fz = Fz + last[2 * (size_t)nn + tid];
            }
            collide<L, FM>(f, r, Fx, Fy, Fz);
            static_for<0, L::Q>([&](auto I) { dst[(size_t)I * g.ns + node] = f[I]; });
        };
        if (own_rows) {
            // The moment rows of stencil node n are sums over the populations lane n has
            // pulled (density: every direction with coefficient 1; momentum a: the directions
            // with c_ia != 0 in lattice order, coefficient c_ia, the row padded with zero
            // terms to q), summed as affine_finish sums a row of q terms: the same bits.
            auto rows_hook = [&](T* moments) {
                if (!has) return;
                auto row = [&](const T (&t)[L::Q]) {
                    T acc = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
#pragma unroll
                    for (int k = 8; k < L::Q; ++k) acc = acc + t[k];
                    return T(0) + acc;
                };
                static_assert(L::Q == 9, "rows of nine terms: one chunk of eight, then one term");
                T t[L::Q];
                static_for<0, L::Q>([&](auto I) { t[I] = T(1) * f[I]; });
                moments[tid] = row(t);
                static_for<0, L::D>([&](auto A) {
                    constexpr int a = A;
                    int used = 0;
                    static_for<0, L::Q>([&](auto I) {
                        constexpr int i = I;
                        constexpr int c = a == 0 ? L::cx[i] : (a == 1 ? L::cy[i] : L::cz[i]);
                        if constexpr (c != 0) t[used++] = T(c) * f[i];
                    });
#pragma unroll
                    for (int k = 0; k < L::Q; ++k)
                        if (k >= used) t[k] = T(0);
                    moments[(size_t)(1 + a) * nn + tid] = row(t);
                });
            };
            ibm_force_group<T, L::D, true, RECORD, decltype(before), decltype(after), true>(
                src, slot, coeff, constant, m, nodes, nn, inverse_g, weights_g, areas_g, nm, start_g, entry_g,
                iterations, hfx, hfy, hfz, g.n, field, body_force, nullptr, record, tid, lanes, before, after, rows_hook);
            return;
        }
        ibm_force_group<T, L::D, true, RECORD>(src, slot, coeff, constant, m, nodes, nn, inverse_g, weights_g, areas_g,
                                               nm, start_g, entry_g, iterations, hfx, hfy, hfz, g.n, field, body_force,
                                               nullptr, record, tid, lanes, before, after);
        return;
    }
    const int chunks = rules.count ? rule_chunk_count(rules.count, rules.chunk) : 0;
    if (b - 1 < chunks) {
        extern __shared__ double dynamic_lds[];
        rule_chunk<L, T, FM>(src, dst, kind, wall_u, field, body, g, r, rules, b - 1, lanes, tid,
                             reinterpret_cast<T*>(dynamic_lds));
        return;
    }
    const int tile = b - 1 - chunks;
    Coord c;
    c.x = (tile % gx) * (int)blockDim.x + (int)threadIdx.x;
    c.y = ((tile / gx) % gy) * (int)blockDim.y + (int)threadIdx.y;
    c.z = tile / (gx * gy);
    if (c.x >= g.nx || c.y >= g.ny) return;
    pull_update<L, T, FM, true, true, true>(src, dst, kind, wall_u, field, body, g, r, c);
}

// ------------------------------------------------------------------ many workgroups
//
// The same forcing in 2 + 2 x iterations launches of many workgroups, for bodies whose
// intermediates do not fit one workgroup's LDS (3D spheres): the launch boundaries take
// the place of the one-workgroup kernel's barriers. Every value is computed by the same
// expression in the same order as k_ibm computes it (see lbm_ibm_core.h), so the two give
// the same bits; only which lane computes it changes. Intermediates live in `scratch` in
// the non-recording layout of ibm_forward_layout (one iteration slot, force updated in
// place), followed by the d running marker totals; u holds the velocity of the next
// iteration.

// One wavefront per workgroup: the work items' gathers are scattered, so spreading them
// over as many compute units as possible shortens each launch (3D sphere, 491 markers,
// about 2,500 stencil nodes: 256-lane groups left the spread on 10 compute units).
constexpr int IBM_MULTI_THREADS = 64;
constexpr int IBM_NODES_PER_GROUP = 16;   // k_ibm_nodes: (D + 1) x 16 moment rows per group
constexpr int IBM_TOTAL_CHUNK = 1024;     // markers staged in LDS per pass of the total (16 per lane)
constexpr int IBM_TOTAL_BATCH = 32;       // staged values read ahead of their sequential adds

// P0 and P1: the D + 1 moment rows of 16 stencil nodes, one row per lane (rows of one
// moment for consecutive nodes are consecutive, so the coefficient loads coalesce), then,
// after the group barrier, per node rho, u*, force = 0 and the velocity of the first
// iteration u = u* + force / (2 rho).
template <class T, int D>
__global__ __launch_bounds__(IBM_MULTI_THREADS) void k_ibm_nodes(const T* __restrict__ F,
                                                                 const int64_t* __restrict__ slot,
                                                                 const T* __restrict__ coeff,
                                                                 const T* __restrict__ constant, int m, int nn, T hfx,
                                                                 T hfy, T hfz, IbmForward<T> s)
{
    static_assert((D + 1) * IBM_NODES_PER_GROUP <= IBM_MULTI_THREADS, "one row per lane");
    __shared__ T mom[D + 1][IBM_NODES_PER_GROUP];
    const int tid = threadIdx.x;
    const int n0 = blockIdx.x * IBM_NODES_PER_GROUP;
    const int moment = tid / IBM_NODES_PER_GROUP;
    const int n = n0 + tid % IBM_NODES_PER_GROUP;
    const int rows = (D + 1) * nn;
    if (moment <= D && n < nn) {
        const int64_t r = (int64_t)moment * nn + n;
        const T v = m <= 24 ? affine_row_wide<T, 24>(F, slot, coeff, constant, rows, m, r)
                            : affine_row(F, slot, coeff, constant, rows, m, r);
        s.moments[r] = v;
        mom[moment][tid % IBM_NODES_PER_GROUP] = v;
    }
    __syncthreads();
    const int node = n0 + tid;
    if (tid < IBM_NODES_PER_GROUP && node < nn) {
        const T half_force[3] = {hfx, hfy, hfz};
        const T rho = T(1) + mom[0][tid];
        s.rho[node] = rho;
        const T two_rho = T(2) * rho;
        for (int a = 0; a < D; ++a) {
            const T ustar = (mom[1 + a][tid] + half_force[a]) / rho;
            s.ustar[(size_t)a * nn + node] = ustar;
            s.force[(size_t)a * nn + node] = T(0);
            s.u[(size_t)a * nn + node] = ustar + T(0) / two_rho;
        }
    }
}

// P2 (first iteration only) and P4 for one marker per lane: rho_l = pairwise sum of
// w rho over the stencil, U_l = sequential sum of w u, F_l = 2 rho_l (0 - U_l) dA_l. The
// stencil's gathers of all velocity components go out together.
template <class T, int D>
__global__ __launch_bounds__(IBM_MULTI_THREADS) void k_ibm_markers(IbmForward<T> s, const int* __restrict__ inverse,
                                                                   const T* __restrict__ weights,
                                                                   const T* __restrict__ areas, int nn, int nm,
                                                                   int first)
{
    constexpr int S = D == 2 ? 9 : 27;
    const int l = blockIdx.x * IBM_MULTI_THREADS + threadIdx.x;
    if (l >= nm) return;
    int inv[S];
    T w[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
        inv[p] = inverse[(size_t)l * S + p];
        w[p] = weights[(size_t)l * S + p];
    }
    hold_loaded(inv);
    T rho_m;
    if (first) {
        T rv[S];
#pragma unroll
        for (int p = 0; p < S; ++p) rv[p] = s.rho[inv[p]];
        hold_loaded(rv);
        rho_m = pairwise_block<T>(S, [&](int p) { return w[p] * rv[p]; });
        s.rho_m[l] = rho_m;
    } else {
        rho_m = s.rho_m[l];
    }
    const T two_rho_m = T(2) * rho_m;
    const T area = areas[l];
    T v[D][S];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int p = 0; p < S; ++p) v[a][p] = s.u[(size_t)a * nn + inv[p]];
#pragma unroll
    for (int a = 0; a < D; ++a) hold_loaded(v[a]);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        T um = w[0] * v[a][0];
#pragma unroll
        for (int p = 1; p < S; ++p) um = um + w[p] * v[a][p];
        s.U[(size_t)a * nm + l] = um;
        s.Fm[(size_t)a * nm + l] = (two_rho_m * (T(0) - um)) * area;
    }
}

// P5 for one stencil node per lane: force += the float64 sum of w_e F_l(e) over the node's
// entries in increasing flat (marker, stencil) order (fetched eight entries at a time), then
// the velocity of the next iteration, or after the last iteration the field scatter. The
// extra last workgroup adds this iteration's marker forces to the totals, sequentially over
// the markers staged through LDS, and after the last iteration writes body_force = -total.
template <class T, int D>
__global__ __launch_bounds__(IBM_MULTI_THREADS) void k_ibm_spread(IbmForward<T> s, IbmTables<T> tb,
                                                                  const int64_t* __restrict__ nodes, int nn, int nm,
                                                                  int first, int last, uint32_t n_total,
                                                                  T* __restrict__ field, T* __restrict__ total,
                                                                  T* __restrict__ body_force)
{
    constexpr int S = D == 2 ? 9 : 27;
    const int tid = threadIdx.x;
    const int node_groups = (nn + IBM_MULTI_THREADS - 1) / IBM_MULTI_THREADS;
    if ((int)blockIdx.x == node_groups) {
        // Staged one component row at a time, so the staging indices need no division; the
        // sequential sum reads its terms from LDS 32 at a time (whole batches carry no bounds
        // tests, so their reads go out together) and adds them one by one.
        __shared__ T stage[D][IBM_TOTAL_CHUNK];
        T sum = T(0);
        for (int c0 = 0; c0 < nm; c0 += IBM_TOTAL_CHUNK) {
            const int count = min(IBM_TOTAL_CHUNK, nm - c0);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                constexpr int AHEAD = IBM_TOTAL_CHUNK / IBM_MULTI_THREADS;
                T v[AHEAD];
#pragma unroll
                for (int k = 0; k < AHEAD; ++k) {
                    const int i = k * IBM_MULTI_THREADS + tid;
                    if (i < count) v[k] = s.Fm[(size_t)a * nm + c0 + i];
                }
#pragma unroll
                for (int k = 0; k < AHEAD; ++k) {
                    const int i = k * IBM_MULTI_THREADS + tid;
                    if (i < count) stage[a][i] = v[k];
                }
            }
            __syncthreads();
            if (tid < D) {
                const T* row = stage[tid];
                int p = 0;
                if (c0 == 0) {  // the sum starts from the first marker's force
                    sum = row[0];
                    p = 1;
                }
                for (; p + IBM_TOTAL_BATCH <= count; p += IBM_TOTAL_BATCH) {
                    T v[IBM_TOTAL_BATCH];
#pragma unroll
                    for (int k = 0; k < IBM_TOTAL_BATCH; ++k) v[k] = row[p + k];
#pragma unroll
                    for (int k = 0; k < IBM_TOTAL_BATCH; ++k) sum = sum + v[k];
                }
                for (; p < count; ++p) sum = sum + row[p];
            }
            __syncthreads();
        }
        if (tid < D) {
            const T t = (first ? T(0) : total[tid]) + sum;
            total[tid] = t;
            if (last) body_force[tid] = -t;
        }
        return;
    }
    const int n = blockIdx.x * IBM_MULTI_THREADS + tid;
    if (n >= nn) return;
    const int begin = tb.spread_start[n], end = tb.spread_start[n + 1];
    double acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0;
    for (int e0 = begin; e0 < end; e0 += 8) {
        int k[8];
        T w[8], f[8][D];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (e0 + j < end) k[j] = tb.spread_entry[e0 + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (e0 + j < end) {
                w[j] = tb.weights[k[j]];
#pragma unroll
                for (int a = 0; a < D; ++a) f[j][a] = s.Fm[(size_t)a * nm + k[j] / S];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (e0 + j < end)
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += (double)(w[j] * f[j][a]);
    }
    const T two_rho = T(2) * s.rho[n];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const T fn = T((double)s.force[(size_t)a * nn + n] + acc[a]);
        s.force[(size_t)a * nn + n] = fn;
        if (last) field[(size_t)a * n_total + nodes[n]] = fn;
        else s.u[(size_t)a * nn + n] = s.ustar[(size_t)a * nn + n] + fn / two_rho;
    }
}

// The launches of the many-workgroup forcing (forward only, no record).
template <class T, int D>
void ibm_force_multi(const T* f, const int64_t* slot, const T* coeff, const T* constant, int m, const int64_t* nodes,
                     int nn, const int* inverse, const T* weights, const T* areas, int nm, const int* start,
                     const int* entry, int iterations, const T (&hf)[3], uint32_t n_total, T* field, T* body_force,
                     T* scratch, hipStream_t st)
{
    IbmForward<T> s;
    T* total = scratch + ibm_forward_layout(scratch, D, nn, nm, 1, &s);
    const IbmTables<T> tb{inverse, weights, areas, start, entry};
    const unsigned node_groups = (unsigned)((nn + IBM_MULTI_THREADS - 1) / IBM_MULTI_THREADS);
    k_ibm_nodes<T, D><<<(nn + IBM_NODES_PER_GROUP - 1) / IBM_NODES_PER_GROUP, IBM_MULTI_THREADS, 0, st>>>(
        f, slot, coeff, constant, m, nn, hf[0], hf[1], hf[2], s);
    for (int it = 0; it < iterations; ++it) {
        k_ibm_markers<T, D><<<(nm + IBM_MULTI_THREADS - 1) / IBM_MULTI_THREADS, IBM_MULTI_THREADS, 0, st>>>(
            s, inverse, weights, areas, nn, nm, it == 0);
        k_ibm_spread<T, D><<<node_groups + 1, IBM_MULTI_THREADS, 0, st>>>(
            s, tb, nodes, nn, nm, it == 0, it == iterations - 1, n_total, field, total, body_force);
    }
}

// Whether k_ibm_step takes the forcing: D2Q9, bx * by in 64 .. 1024 lanes, at least one lane
// per stencil node, and the staged forcing (with `record`'s slots) in the LDS.
bool ibm_step_ok(int lattice, int dtype, int n_nodes, int n_markers, int iterations, int record, int bx, int by)
{
    if (lattice != LBM_D2Q9 || (dtype != LBM_F32 && dtype != LBM_F64) || bx <= 0 || by <= 0) return false;
    const int lanes = bx * by;
    if (lanes < RULE_CHUNK_MAX || lanes > 1024 || n_nodes <= 0 || n_nodes > lanes || n_markers <= 0 ||
        iterations < 1)
        return false;
    const size_t real = dtype == LBM_F32 ? sizeof(float) : sizeof(double);
    return ibm_lds_bytes(2, n_nodes, n_markers, record ? iterations : 1, real) <= lds_limit();
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LBM_API int lbm_ibm_pull_step_supported(int lattice, int dtype, int n_nodes, int n_markers, int iterations,
                                        int record, int bx, int by, int* ok)
{
    return guarded("lbm_ibm_pull_step_supported", [&] {
        require(ok != nullptr, "null output");
        *ok = ibm_step_ok(lattice, dtype, n_nodes, n_markers, iterations, record, bx, by) ? 1 : 0;
    });
}

LBM_API int lbm_ibm_pull_step(int lattice, int dtype, const void* src, void* dst, const uint8_t* kind,
                              const void* wall, void* field, const void* body, int nx, int ny, int nz,
                              double omega_plus, double omega_minus, double fx, double fy, double fz, int force_mode,
                              const int64_t* slot, const void* coeff, const void* constant, int m,
                              const int64_t* nodes, int n_nodes, const int* inverse, const void* weights,
                              const void* areas, int n_markers, const int* spread_start, const int* spread_entry,
                              int iterations, void* body_force, void* record, const uint32_t* rule_node_info,
                              const int32_t* rule_node_first, int rule_nodes, const int64_t* rule_slot,
                              const void* rule_coeff, const void* rule_constant, int64_t n_rules, int rule_m,
                              int rows_in_lanes, int bx, int by, void* stream)
{
    return guarded("lbm_ibm_pull_step", [&] {
        require(ibm_step_ok(lattice, dtype, n_nodes, n_markers, iterations, record != nullptr, bx, by),
                "the forcing inside the step takes D2Q9, 64 .. 1024 lanes, a lane per stencil node and a forcing "
                "that fits the LDS (lbm_ibm_pull_step_supported)");
        const Grid g = make_grid(lattice, nx, ny, nz);
        require(src && dst && src != dst && kind && field && body_force, "null or aliased step buffers");
        require(force_mode == 1 || force_mode == 2, "the forced step uses force mode field (1) or dense (2)");
        require(force_mode != 2 || body != nullptr, "dense force mode needs a body field");
        require(m >= 1 && m <= 128, "moment rows need 1 <= m <= 128 terms");
        require(slot && coeff && constant && nodes && inverse && weights && areas && spread_start && spread_entry,
                "null IBM table");
        require(n_rules >= 0 && n_rules < (1ll << 31), "rule count out of range");
        require(rule_nodes >= 0 && (n_rules > 0) == (rule_nodes > 0), "rows need rule nodes and nodes rows");
        require(n_rules == 0 || (rule_node_info && rule_node_first && rule_slot && rule_coeff && rule_constant &&
                                 rule_m >= 1 && rule_m <= 128),
                "link rules need the node list, slot, coeff, constant and 1 <= m <= 128");
        require(reinterpret_cast<uintptr_t>(rule_node_info) % 16 == 0, "rule_node_info must be 16-byte aligned");
        require(!rows_in_lanes || m == (lattice == LBM_D2Q9 ? 9 : 19), "rows_in_lanes needs rows of q terms");
        const double force[3] = {fx, fy, fz};
        const dim3 block((unsigned)bx, (unsigned)by, 1);
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            if constexpr (L::D == 2) {
                const Relaxation<T> r = make_relaxation<L, T>(omega_plus, omega_minus, force);
                const dim3 tiles = grid_of(g, block);
                const int64_t n_tiles = (int64_t)tiles.x * tiles.y * tiles.z;
                const int chunk = rule_chunk_nodes(rule_nodes, (int)std::min<int64_t>(n_tiles, 1 << 20));
                const int chunks = rule_nodes ? rule_chunk_count(rule_nodes, chunk) : 0;
                const RuleRows<T> rr{reinterpret_cast<const uint4*>(rule_node_info), rule_node_first, rule_nodes,
                                     chunk, rule_slot, static_cast<const T*>(rule_coeff),
                                     static_cast<const T*>(rule_constant), n_rules, rule_m};
                require(1 + chunks + n_tiles < (1ll << 31), "grid too large for a one-dimensional launch");
                const size_t lds =
                    std::max(ibm_lds_bytes(L::D, n_nodes, n_markers, record ? iterations : 1, sizeof(T)),
                             chunks ? sizeof(T) * (size_t)chunk * (L::Q - 1) : (size_t)0);
                auto go = [&](auto fm, auto recorded) {
                    constexpr ForceMode FM = decltype(fm)::value;
                    constexpr bool RE = decltype(recorded)::value;
                    allow_lds(k_ibm_step<L, T, FM, RE>, lds);
                    k_ibm_step<L, T, FM, RE><<<(unsigned)(1 + chunks + n_tiles), block, lds, as_stream(stream)>>>(
                        static_cast<const T*>(src), slot, static_cast<const T*>(coeff),
                        static_cast<const T*>(constant), m, nodes, n_nodes, inverse, static_cast<const T*>(weights),
                        static_cast<const T*>(areas), n_markers, spread_start, spread_entry, iterations,
                        T(0.5 * fx), T(0.5 * fy), T(0.5 * fz), static_cast<T*>(field), static_cast<T*>(body_force),
                        static_cast<T*>(record), static_cast<T*>(dst), kind, static_cast<const T*>(wall),
                        static_cast<const T*>(body), g, r, rr, (int)tiles.x, (int)tiles.y, rows_in_lanes);
                };
                using Fi = std::integral_constant<ForceMode, ForceMode::FIELD>;
                using De = std::integral_constant<ForceMode, ForceMode::DENSE>;
                if (force_mode == 1 && record) go(Fi{}, std::true_type{});
                else if (force_mode == 1) go(Fi{}, std::false_type{});
                else if (record) go(De{}, std::true_type{});
                else go(De{}, std::false_type{});
            } else {
                require(false, "the forcing inside the step takes D2Q9");
            }
            check_launch("IBM forcing inside the pull step");
        });
    });
}

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
            if (!staged && !record) {
                ibm_force_multi<T, L::D>(static_cast<const T*>(f), slot, static_cast<const T*>(coeff),
                                         static_cast<const T*>(constant), m, nodes, n_nodes, inverse,
                                         static_cast<const T*>(weights), static_cast<const T*>(areas), n_markers,
                                         spread_start, spread_entry, iterations, hf, (uint32_t)n_total,
                                         static_cast<T*>(field), static_cast<T*>(body_force),
                                         static_cast<T*>(scratch), as_stream(stream));
                check_launch("IBM kernels");
                return;
            }
            if (staged && record) go(std::true_type{}, std::true_type{});
            else if (staged) go(std::true_type{}, std::false_type{});
            else if (record) go(std::false_type{}, std::true_type{});
            else go(std::false_type{}, std::false_type{});
            check_launch("IBM kernel");
        });
    });
}

LBM_API int lbm_ibm_force_multi(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                                const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                                const void* weights, const void* areas, int n_markers, const int* spread_start,
                                const int* spread_entry, int iterations, double fx, double fy, double fz,
                                int64_t n_total, void* field, void* body_force, void* scratch, void* stream)
{
    return guarded("lbm_ibm_force_multi", [&] {
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
            ibm_force_multi<T, L::D>(static_cast<const T*>(f), slot, static_cast<const T*>(coeff),
                                     static_cast<const T*>(constant), m, nodes, n_nodes, inverse,
                                     static_cast<const T*>(weights), static_cast<const T*>(areas), n_markers,
                                     spread_start, spread_entry, iterations, hf, (uint32_t)n_total,
                                     static_cast<T*>(field), static_cast<T*>(body_force), static_cast<T*>(scratch),
                                     as_stream(stream));
            check_launch("IBM kernels");
        });
    });
}

}  // extern "C"
