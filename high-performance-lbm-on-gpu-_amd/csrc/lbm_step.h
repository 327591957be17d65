// Device building blocks shared by the stream-collide kernels: periodic neighbour
// indexing, near-wall masks, per-node forces, and the fused pull update of one node.
#pragma once

#include "lbm_common.h"
#include "lbm_hold.h"

namespace lbm {

// ------------------------------------------------------------------ indexing

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

template <class L, int I, int S>
__device__ __forceinline__ uint32_t neighbour(int x, int y, int z, const Grid& g)
{
    using Di = Dir<L, I>;
    const int xx = Di::x == 0 ? x : wrap(x + S * Di::x, g.nx);
    const int yy = Di::y == 0 ? y : wrap(y + S * Di::y, g.ny);
    const int zz = (L::D == 2 || Di::z == 0) ? z : wrap(z + S * Di::z, g.nz);
    return ((uint32_t)zz * (uint32_t)g.ny + (uint32_t)yy) * (uint32_t)g.nx + (uint32_t)xx;
}

// Bit i set when the upstream node x - c_i is solid. The Q - 1 neighbour kinds are read
// together and held, then tested (else each read could sit behind its own test and wait).
template <class L>
__device__ __forceinline__ uint32_t blocked_upstream(const uint8_t* __restrict__ kind, int x, int y, int z,
                                                     const Grid& g)
{
    uint32_t nk[L::Q];
    nk[0] = 0;
    static_for<1, L::Q>([&](auto I) { nk[I] = kind[neighbour<L, I, -1>(x, y, z, g)]; });
    hold_loaded(nk);
    uint32_t m = 0;
    static_for<1, L::Q>([&](auto I) {
        if (nk[I] & KIND_SOLID) m |= 1u << I;
    });
    return m;
}

// Force of one node for the non-uniform modes.
template <class L, ForceMode FM, class T>
__device__ __forceinline__ void node_force(const Relaxation<T>& r, uint8_t k, const T* __restrict__ field,
                                           const T* __restrict__ body, uint32_t node, uint32_t n, T& Fx,
                                           T& Fy, T& Fz)
{
    Fx = r.force[0];
    Fy = r.force[1];
    Fz = r.force[2];
    if constexpr (FM == ForceMode::DENSE) {
        Fx = Fx + body[node];
        Fy = Fy + body[(size_t)n + node];
        if constexpr (L::D == 3) Fz = Fz + body[2 * (size_t)n + node];
    }
    if constexpr (FM != ForceMode::UNIFORM) {
        if (k & KIND_FORCED) {
            Fx = Fx + field[node];
            Fy = Fy + field[(size_t)n + node];
            if constexpr (L::D == 3) Fz = Fz + field[2 * (size_t)n + node];
        }
    }
}

struct Coord {
    int x, y, z;
};

// f_i += the moving-wall term of link i for every set bit i of `bits`, with u_w read at
// x + S c_i. All wall-velocity loads of the node issue together (the links' terms are
// evaluated unconditionally, then selected), so a near-wall wavefront waits once. A null
// wall array means every wall is at rest: each term is an exact zero and nothing is read.
template <class L, int S, class T>
__device__ __forceinline__ void add_wall_terms(T (&f)[L::Q], uint32_t bits, const T* __restrict__ wall_u,
                                               const Coord& c, const Grid& g)
{
    if (wall_u == nullptr) return;
    T w[L::Q];
    static_for<0, L::Q>([&](auto I) { w[I] = wall_term<L, I>(wall_u, neighbour<L, I, S>(c.x, c.y, c.z, g), g.n); });
    hold_loaded(w);
    static_for<0, L::Q>([&](auto I) {
        if ((bits >> I) & 1u) f[I] = f[I] + w[I];
    });
}

__device__ __forceinline__ bool locate(Coord& c, const Grid& g)
{
    c.x = blockIdx.x * blockDim.x + threadIdx.x;
    c.y = blockIdx.y * blockDim.y + threadIdx.y;
    c.z = blockIdx.z;
    return c.x < g.nx && c.y < g.ny;
}

__device__ __forceinline__ uint32_t flat(const Coord& c, const Grid& g)
{
    return ((uint32_t)c.z * (uint32_t)g.ny + (uint32_t)c.y) * (uint32_t)g.nx + (uint32_t)c.x;
}

// ------------------------------------------------------------------ kernels

// One node of the two-lattice fused pull step (shared by the single-lattice and the
// batched kernels, so both produce the same bits). Population planes are g.ns elements
// apart (n for one lattice; padded in the batched pools); wall and force fields g.n.
// SKIP_RULES: nodes whose kind has KIND_RULES are left to the rule layer of the step
// (k_pull_links), which updates them with their link rules. SKIP_REGION: nodes whose kind
// has KIND_ADJ_REGION are left to a region workgroup of a fused step.
template <class L, class T, ForceMode FM, bool MASK, bool SKIP_RULES = false, bool SKIP_REGION = false>
__device__ __forceinline__ void pull_update(const T* __restrict__ src, T* __restrict__ dst,
                                            const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                            const T* __restrict__ field, const T* __restrict__ body, const Grid& g,
                                            const Relaxation<T>& r, const Coord& c)
{
    static_assert(MASK || !(SKIP_RULES || SKIP_REGION), "the skipped nodes are marked in the kind bytes");
    constexpr uint8_t skip =
        KIND_SOLID | (SKIP_RULES ? KIND_RULES : uint8_t(0)) | (SKIP_REGION ? KIND_ADJ_REGION : uint8_t(0));
    const uint32_t node = flat(c, g);
    // The node's kind byte and all Q pulled populations are loaded together (the pulled
    // slots are valid addresses whatever the node is; a solid node discards them). Only a
    // node with a solid upstream neighbour waits for the neighbour kinds, then reloads
    // each population from its selected slot (bounce-back or pulled) and adds the
    // moving-wall terms, all Q loads of each pass in flight together.
    const uint8_t k = MASK ? kind[node] : 0;
    T f[L::Q];
    static_for<0, L::Q>([&](auto I) { f[I] = src[(size_t)I * g.ns + neighbour<L, I, -1>(c.x, c.y, c.z, g)]; });
    if constexpr (MASK) hold_loaded(f);
    if (MASK && (k & skip)) return;
    if (MASK && (k & KIND_NEAR_SOLID)) {
        const uint32_t blocked = blocked_upstream<L>(kind, c.x, c.y, c.z, g);
        if (blocked) {
            static_for<0, L::Q>([&](auto I) {
                constexpr int i = I;
                const bool bounce = (blocked >> i) & 1u;
                f[i] = src[bounce ? (size_t)L::opp[i] * g.ns + node
                                  : (size_t)i * g.ns + neighbour<L, i, -1>(c.x, c.y, c.z, g)];
            });
            add_wall_terms<L, -1>(f, blocked, wall_u, c, g);
        }
    }
    T Fx = T(0), Fy = T(0), Fz = T(0);
    if constexpr (FM != ForceMode::UNIFORM) node_force<L, FM>(r, k, field, body, node, g.n, Fx, Fy, Fz);
    collide<L, FM>(f, r, Fx, Fy, Fz);
    static_for<0, L::Q>([&](auto I) { dst[(size_t)I * g.ns + node] = f[I]; });
}

}  // namespace lbm
