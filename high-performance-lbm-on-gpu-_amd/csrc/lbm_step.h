// Device building blocks shared by the stream-collide kernels: periodic neighbour
// indexing, near-wall masks, per-node forces, and the fused pull update of one node.
#pragma once

#include "lbm_common.h"

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

// Bit i set when the upstream node x - c_i is solid.
template <class L>
__device__ __forceinline__ uint32_t blocked_upstream(const uint8_t* __restrict__ kind, int x, int y, int z,
                                                     const Grid& g)
{
    uint32_t m = 0;
    static_for<1, L::Q>([&](auto I) {
        if (kind[neighbour<L, I, -1>(x, y, z, g)] & KIND_SOLID) m |= 1u << I;
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
// batched kernels, so both produce the same bits).
template <class L, class T, ForceMode FM, bool MASK>
__device__ __forceinline__ void pull_update(const T* __restrict__ src, T* __restrict__ dst,
                                            const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                            const T* __restrict__ field, const T* __restrict__ body, const Grid& g,
                                            const Relaxation<T>& r, const Coord& c)
{
    const uint32_t node = flat(c, g);
    const uint8_t k = MASK ? kind[node] : 0;
    if (MASK && (k & KIND_SOLID)) return;
    const uint32_t blocked = (MASK && (k & KIND_NEAR_SOLID)) ? blocked_upstream<L>(kind, c.x, c.y, c.z, g) : 0u;

    T f[L::Q];
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        const uint32_t up = neighbour<L, i, -1>(c.x, c.y, c.z, g);
        if (MASK && ((blocked >> i) & 1u)) {
            f[i] = src[(size_t)L::opp[i] * g.n + node] + wall_term<L, i>(wall_u, up, g.n);
        } else {
            f[i] = src[(size_t)i * g.n + up];
        }
    });
    T Fx = T(0), Fy = T(0), Fz = T(0);
    if constexpr (FM != ForceMode::UNIFORM) node_force<L, FM>(r, k, field, body, node, g.n, Fx, Fy, Fz);
    collide<L, FM>(f, r, Fx, Fy, Fz);
    static_for<0, L::Q>([&](auto I) { dst[(size_t)I * g.n + node] = f[I]; });
}

}  // namespace lbm
