// Stream-collide kernels (two-lattice pull, AA even/odd, LDS-tiled), initialisation,
// macroscopic readout and the divergence watchdog, with their C-ABI entry points.
//
// Thread mapping: one thread per node, blockDim (bx, by), grid (ceil(nx/bx),
// ceil(ny/by), nz). x is the fastest axis of every population plane, so the 64
// lanes of a wavefront touch 64 consecutive addresses of each plane; with
// bx = nx (whole rows per workgroup) every cache line of a plane is read or
// written by exactly one workgroup per step, shifted accesses included.
// Plane base pointers (f + i * n) are wave-uniform, per-lane offsets 32-bit.
//
// Memory traffic per fluid node update is Q loads + Q stores of real_t
// (152 B for D3Q19 float32) in every variant; the kind byte adds 1 B.

#include "lbm_common.h"
#include "lbm_force.h"
#include "lbm_rule_layer.h"
#include "lbm_step.h"

#include <cstring>
#include <mutex>
#include <string>

namespace lbm {

// Two-lattice fused pull step.
template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(step_max_threads(L::D, sizeof(T))) void k_pull(const T* __restrict__ src, T* __restrict__ dst,
                                               const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                               const T* __restrict__ field, const T* __restrict__ body, Grid g,
                                               Relaxation<T> r)
{
    Coord c;
    if (!locate(c, g)) return;
    pull_update<L, T, FM, MASK>(src, dst, kind, wall_u, field, body, g, r, c);
}

// ------------------------------------------------------------------ pull step with link rules
// (the rule layer, RuleRows and rule_chunk, is in lbm_rule_layer.h)

// Two-lattice pull step with its link rules and, optionally, the momentum-exchange force of
// the state it reads. The grid's z layers 0..nz-1 update every node except solid and rule
// nodes (pull_update); the extra layer z = nz hosts, in turn, the rule chunks (rule_chunk)
// and then the force probe's reduction groups (force_group_hosted, the groups and sums of
// lbm_affine_force). Nothing in the launch writes src, so the force rows read the state
// exactly as a separate force launch before the step would. The dynamic LDS holds the rule
// values of a chunk (lanes values) or the force tree (3 x FORCE_WAVES doubles).
template <class L, class T, ForceMode FM>
__global__ __launch_bounds__(step_max_threads(L::D, sizeof(T))) void k_pull_links(
    const T* __restrict__ src, T* __restrict__ dst, const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
    const T* __restrict__ field, const T* __restrict__ body, Grid g, Relaxation<T> r, RuleRows<T> rules,
    ForceRows<T> fr, int force_groups, double* __restrict__ force_out, double* __restrict__ force_partial, int gx,
    int gy)
{
    // A one-dimensional grid: the first `work` workgroups take the rule chunks and the force
    // groups (so they start first), the rest the (gx, gy, nz) tiles of the lattice.
    const int chunks = rules.count ? rule_chunk_count(rules.count, rules.chunk) : 0;
    const int work = chunks + force_groups;
    const int b = (int)blockIdx.x;
    if (b < work) {
        extern __shared__ double dynamic_lds[];
        const int lanes = blockDim.x * blockDim.y;
        const int tid = threadIdx.y * blockDim.x + threadIdx.x;
        if (b < chunks) {
            rule_chunk<L, T, FM>(src, dst, kind, wall_u, field, body, g, r, rules, b, lanes, tid,
                                 reinterpret_cast<T*>(dynamic_lds));
        } else {
            auto& part = *reinterpret_cast<double(*)[3][FORCE_WAVES]>(dynamic_lds);
            force_group_hosted(src, fr, force_groups, b - chunks, lanes, tid, part, force_out, force_partial);
        }
        return;
    }
    const int tile = b - work;
    Coord c;
    c.x = (tile % gx) * (int)blockDim.x + (int)threadIdx.x;
    c.y = ((tile / gx) % gy) * (int)blockDim.y + (int)threadIdx.y;
    c.z = tile / (gx * gy);
    if (c.x >= g.nx || c.y >= g.ny) return;
    pull_update<L, T, FM, true, true>(src, dst, kind, wall_u, field, body, g, r, c);
}

// AA even half-step: node-local, slot reversal.
template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(step_max_threads(L::D, sizeof(T))) void k_aa_even(T* __restrict__ F, const uint8_t* __restrict__ kind,
                                                  const T* __restrict__ field, const T* __restrict__ body,
                                                  Grid g, Relaxation<T> r)
{
    Coord c;
    if (!locate(c, g)) return;
    const uint32_t node = flat(c, g);
    // The kind byte and the Q populations are loaded together; a solid node discards them.
    const uint8_t k = MASK ? kind[node] : 0;
    T f[L::Q];
    static_for<0, L::Q>([&](auto I) { f[I] = F[(size_t)I * g.n + node]; });
    if constexpr (MASK) hold_loaded(f);
    if (MASK && (k & KIND_SOLID)) return;
    T Fx = T(0), Fy = T(0), Fz = T(0);
    if constexpr (FM != ForceMode::UNIFORM) node_force<L, FM>(r, k, field, body, node, g.n, Fx, Fy, Fz);
    collide<L, FM>(f, r, Fx, Fy, Fz);
    static_for<0, L::Q>([&](auto I) { F[(size_t)L::opp[I] * g.n + node] = f[I]; });
}

// AA odd half-step: pull from x - c_i, push to x + c_i, in place. Slot (k, y) is
// read and written by the same thread only (src/lbm/backends/kernels/stream_collide_aa.cu:15-31),
// and every thread loads all Q values before it stores any.
template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(step_max_threads(L::D, sizeof(T))) void k_aa_odd(T* __restrict__ F, const uint8_t* __restrict__ kind,
                                                 const T* __restrict__ wall_u, const T* __restrict__ field,
                                                 const T* __restrict__ body, Grid g, Relaxation<T> r)
{
    Coord c;
    if (!locate(c, g)) return;
    const uint32_t node = flat(c, g);
    // The kind byte and the Q pulled populations are loaded together (a solid node
    // discards them; it writes nothing). A node with a solid upstream neighbour then
    // reloads each population from its selected slot (bounce-back or pulled) and adds the
    // moving-wall terms. Loads and stores go to selected slots rather than branches, so
    // all Q of each pass stay in flight together.
    const uint8_t k = MASK ? kind[node] : 0;
    T f[L::Q];
    static_for<0, L::Q>([&](auto I) {
        f[I] = F[(size_t)L::opp[I] * g.n + neighbour<L, I, -1>(c.x, c.y, c.z, g)];
    });
    if constexpr (MASK) hold_loaded(f);
    if (MASK && (k & KIND_SOLID)) return;
    const uint32_t blocked = (MASK && (k & KIND_NEAR_SOLID)) ? blocked_upstream<L>(kind, c.x, c.y, c.z, g) : 0u;
    // w[i]: moving-wall term of link i, u_w read at the solid node x - c_i (meaningful for
    // the set bits of `blocked` only). It serves twice: the pulled f_i bounces back with
    // +w[i], and after the collision f*_opp(i), which would stream into that same solid node
    // x + c_opp(i) = x - c_i, bounces back into slot i of this node with +w[i] as well (the
    // term of direction i at that node) — so the wall velocities are loaded once.
    // A null wall array means every wall is at rest: the terms are exact zeros, not read.
    const bool moving = MASK && blocked && wall_u != nullptr;
    T w[L::Q];
    if (MASK && blocked) {
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            const bool bounce = (blocked >> i) & 1u;
            f[i] = F[bounce ? (size_t)i * g.n + node
                            : (size_t)L::opp[i] * g.n + neighbour<L, i, -1>(c.x, c.y, c.z, g)];
        });
    }
    if (moving) {
        static_for<0, L::Q>([&](auto I) { w[I] = wall_term<L, I>(wall_u, neighbour<L, I, -1>(c.x, c.y, c.z, g), g.n); });
        static_for<0, L::Q>([&](auto I) {
            if ((blocked >> I) & 1u) f[I] = f[I] + w[I];
        });
    }
    T Fx = T(0), Fy = T(0), Fz = T(0);
    if constexpr (FM != ForceMode::UNIFORM) node_force<L, FM>(r, k, field, body, node, g.n, Fx, Fy, Fz);
    collide<L, FM>(f, r, Fx, Fy, Fz);
    if (moving) {
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            if ((blocked >> i) & 1u) f[L::opp[i]] = f[L::opp[i]] + w[i];
        });
    }
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        constexpr int o = L::opp[i];
        const uint32_t down = neighbour<L, i, +1>(c.x, c.y, c.z, g);
        const bool bounce = MASK && ((blocked >> o) & 1u);
        F[bounce ? (size_t)o * g.n + node : (size_t)i * g.n + down] = f[i];
    });
}

// Periodic image of a tile coordinate v in [-1, n + block): one conditional
// add or subtract, and a modulo only for lanes of a block wider than the grid.
__device__ __forceinline__ int tile_wrap(int v, int n)
{
    if (v < 0) return v + n;
    if (v < n) return v;
    v -= n;
    return v < n ? v : v % n;
}

// Stage one tile cell (ex, ey) in [0, hx) x [0, hy) and its kind byte: Q loads and Q LDS
// stores, unconditional. Of a halo cell (halo) the tile pulls only the planes i whose
// downstream cell (ex + c_ix, ey + c_iy) lies in the tile (3 of 9 on an edge, 1 at a
// corner); an unpulled plane is read at the same row in the nearest column of the tile
// (ix) and lands in an LDS slot that nothing reads. For the halo columns and the corners,
// which belong to the tiles next door in x, that keeps their cache lines out of this
// workgroup's traffic: the round-robin dispatch over the XCDs (each with its own L2) gives
// those tiles to other XCDs, so their lines were fetched from HBM a second time. Cells of
// the halo rows read themselves (ix = ex), as before.
template <class L, class T, bool MASK>
__device__ __forceinline__ void stage_cell(T* tile, uint8_t* kind_tile, int cells, int hx, int hy, int ex, int ey,
                                           bool halo, int x0, int y0, const T* __restrict__ src,
                                           const uint8_t* __restrict__ kind, const Grid& g)
{
    const int gx = tile_wrap(x0 + ex - 1, g.nx), gy = tile_wrap(y0 + ey - 1, g.ny);
    const uint32_t gnode = (uint32_t)gy * (uint32_t)g.nx + (uint32_t)gx;
    const int ix = min(max(ex, 1), hx - 2);
    // the offset to gnode as a value: a select between the two variables themselves is
    // compiled as a select of their addresses, which puts both on the stack
    const uint32_t inner = (uint32_t)gy * (uint32_t)g.nx + (uint32_t)tile_wrap(x0 + ix - 1, g.nx);
    const uint32_t step_out = gnode - inner;
    const int e = ey * hx + ex;
    T v[L::Q];
    static_for<0, L::Q>([&](auto I) {
        using Di = Dir<L, I>;
        const int dx = ex + Di::x, dy = ey + Di::y;
        const bool pulled = !halo || (dx >= 1 && dx < hx - 1 && dy >= 1 && dy < hy - 1);
        v[I] = src[(size_t)I * g.n + (inner + (pulled ? step_out : 0u))];
    });
    kind_tile[e] = MASK ? kind[gnode] : 0;
    static_for<0, L::Q>([&](auto I) { tile[I * cells + e] = v[I]; });
}

// LDS-tiled two-lattice step (2D): the block's tile (all Q planes) plus a one-node halo
// (the planes pulled into the tile) and the kind bytes of both are staged in LDS, then each
// thread pulls its neighbours from LDS. Tile = block; out-of-domain threads help load and
// reach the barrier (src/lbm/backends/kernels/stream_collide_tiled.cu:24-79). Each
// thread stages its own cell (coalesced rows, no index division); the halo ring
// of 2 (bx + 2) + 2 by cells is staged by the first threads in a second pass.
template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(1024) void k_lds(const T* __restrict__ src, T* __restrict__ dst,
                                              const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                              const T* __restrict__ field, const T* __restrict__ body, Grid g,
                                              Relaxation<T> r)
{
    static_assert(L::D == 2, "the LDS-tiled kernel is 2D only");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int bx = blockDim.x, by = blockDim.y;
    const int hx = bx + 2, hy = by + 2;
    const int cells = hx * hy;
    T* tile = reinterpret_cast<T*>(lds_raw);
    uint8_t* kind_tile = lds_raw + sizeof(T) * (size_t)L::Q * cells;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = blockIdx.x * bx, y0 = blockIdx.y * by;
    stage_cell<L, T, MASK>(tile, kind_tile, cells, hx, hy, tx + 1, ty + 1, false, x0, y0, src, kind, g);
    const int ring = 2 * hx + 2 * by;
    for (int h = ty * bx + tx; h < ring; h += bx * by) {
        int ex, ey;
        if (h < 2 * hx) {
            ex = h < hx ? h : h - hx;
            ey = h < hx ? 0 : hy - 1;
        } else {
            const int c = h - 2 * hx;
            ex = c < by ? 0 : hx - 1;
            ey = 1 + (c < by ? c : c - by);
        }
        stage_cell<L, T, MASK>(tile, kind_tile, cells, hx, hy, ex, ey, true, x0, y0, src, kind, g);
    }
    __syncthreads();
    const int x = x0 + tx, y = y0 + ty;
    const int centre = (ty + 1) * hx + (tx + 1);
    const uint8_t k = kind_tile[centre];
    if (x >= g.nx || y >= g.ny || (MASK && (k & KIND_SOLID))) return;
    const uint32_t node = (uint32_t)y * (uint32_t)g.nx + (uint32_t)x;
    T f[L::Q];
    uint32_t blocked = 0;
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        using Di = Dir<L, i>;
        const int e = (ty + 1 - Di::y) * hx + (tx + 1 - Di::x);
        const bool bounce = MASK && (kind_tile[e] & KIND_SOLID);
        if (bounce) blocked |= 1u << i;
        f[i] = tile[bounce ? L::opp[i] * cells + centre : i * cells + e];
    });
    if (MASK && blocked && wall_u != nullptr) {  // null: every wall at rest, zero terms
        // Every link's term from its upstream node (a valid address for any link), held
        // together so the loads go out at once, then added on the blocked links only.
        T w[L::Q];
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            using Di = Dir<L, i>;
            const uint32_t up = (uint32_t)wrap(y - Di::y, g.ny) * (uint32_t)g.nx + (uint32_t)wrap(x - Di::x, g.nx);
            w[i] = wall_term<L, i>(wall_u, up, g.n);
        });
        hold_loaded(w);
        static_for<0, L::Q>([&](auto I) {
            if ((blocked >> I) & 1u) f[I] = f[I] + w[I];
        });
    }
    T Fx = T(0), Fy = T(0), Fz = T(0);
    if constexpr (FM != ForceMode::UNIFORM) node_force<L, FM>(r, k, field, body, node, g.n, Fx, Fy, Fz);
    collide<L, FM>(f, r, Fx, Fy, Fz);
    static_for<0, L::Q>([&](auto I) { dst[(size_t)I * g.n + node] = f[I]; });
}

template <class L, class T>
__global__ __launch_bounds__(256) void k_init(T* __restrict__ F, const T* __restrict__ drho,
                                              const T* __restrict__ ux, const T* __restrict__ uy,
                                              const T* __restrict__ uz, uint32_t n, int aa)
{
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n) return;
    const T dr = drho[node];
    const T vx = ux[node], vy = uy[node];
    const T vz = L::D == 3 ? uz[node] : T(0);
    const T rho = T(1) + dr;
    T usum = vx * vx + vy * vy;
    if constexpr (L::D == 3) usum = usum + vz * vz;
    const T usq = T(1.5) * usum;
    static_for<0, L::Q>([&](auto I) {
        const T feq = shifted_feq<L, I>(dr, rho, vx, vy, vz, usq);
        const int slot = aa ? L::opp[I] : int(I);
        F[(size_t)slot * n + node] = feq;
    });
}

template <class L, class T, ForceMode FM>
__global__ __launch_bounds__(256) void k_macro(const T* __restrict__ F, const uint8_t* __restrict__ kind,
                                               const T* __restrict__ field, uint32_t n, int aa, Relaxation<T> r,
                                               T* __restrict__ rho_out, T* __restrict__ u_out)
{
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n) return;
    const uint8_t k = kind ? kind[node] : 0;
    if (k & KIND_SOLID) {
        rho_out[node] = T(1);
        static_for<0, L::D>([&](auto A) { u_out[(size_t)A * n + node] = T(0); });
        return;
    }
    T f[L::Q];
    static_for<0, L::Q>([&](auto I) { f[I] = F[(size_t)(aa ? L::opp[I] : int(I)) * n + node]; });
    T drho, jx, jy, jz;
    moments<L>(f, drho, jx, jy, jz);
    const T rho = T(1) + drho;
    T h[3] = {r.half_force[0], r.half_force[1], r.half_force[2]};
    if constexpr (FM != ForceMode::UNIFORM) {
        T Fa[3] = {r.force[0], r.force[1], r.force[2]};
        if (k & KIND_FORCED) {
            static_for<0, L::D>([&](auto A) { Fa[A] = Fa[A] + field[(size_t)A * n + node]; });
        }
        static_for<0, 3>([&](auto A) { h[A] = T(0.5) * Fa[A]; });
    }
    rho_out[node] = rho;
    u_out[node] = (jx - h[0]) / rho;
    u_out[(size_t)n + node] = (jy - h[1]) / rho;
    if constexpr (L::D == 3) u_out[2 * (size_t)n + node] = (jz - h[2]) / rho;
}

template <class L, class T>
__global__ __launch_bounds__(256) void k_health(const T* __restrict__ F, const uint8_t* __restrict__ kind,
                                                uint32_t n, int aa, uint32_t* __restrict__ out)
{
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bad = 0xffffffffu;
    uint32_t nonfinite = 0;
    if (node < n && !(kind && (kind[node] & KIND_SOLID))) {
        T drho = F[(size_t)(aa ? L::opp[0] : 0) * n + node];
        static_for<1, L::Q>([&](auto I) { drho = drho + F[(size_t)(aa ? L::opp[I] : int(I)) * n + node]; });
        const bool finite = isfinite(drho);
        if (!finite) nonfinite = 1;
        if (!finite || drho <= T(-1)) bad = node;
    }
    // Wavefront reduction first, one atomic per wave.
    for (int offset = warpSize / 2; offset > 0; offset /= 2) {
        bad = min(bad, (uint32_t)__shfl_xor((int)bad, offset));
        nonfinite |= (uint32_t)__shfl_xor((int)nonfinite, offset);
    }
    if ((threadIdx.x % warpSize) == 0) {
        if (bad != 0xffffffffu) atomicMin(out, bad);
        if (nonfinite) atomicOr(out + 1, 1u);
    }
}

template <class L, class T>
__global__ void k_tables(int* c, int* opp, T* w)
{
    const int i = threadIdx.x;
    if (i >= L::Q) return;
    c[i] = L::cx[i];
    c[L::Q + i] = L::cy[i];
    c[2 * L::Q + i] = L::cz[i];
    opp[i] = L::opp[i];
    w[i] = T(L::w[i]);
}

// ------------------------------------------------------------------ launch helpers

enum class Variant { PULL, EVEN, ODD, LDS };

template <class L, class T, ForceMode FM, bool MASK>
void launch_step(Variant v, const void* a, void* b, const uint8_t* kind, const void* wall, const void* field,
                 const void* body, const Grid& g, const Relaxation<T>& r, dim3 block, int groups_per_cu,
                 hipStream_t s)
{
    const dim3 grid = grid_of(g, block);
    const T* wall_t = static_cast<const T*>(wall);
    const T* field_t = static_cast<const T*>(field);
    const T* body_t = static_cast<const T*>(body);
    // The streaming kernels use no LDS; a cap on resident workgroups is requested as LDS.
    const size_t cap = occupancy_lds(groups_per_cu, 0);
    switch (v) {
    case Variant::PULL:
        allow_lds(k_pull<L, T, FM, MASK>, cap);
        k_pull<L, T, FM, MASK><<<grid, block, cap, s>>>(static_cast<const T*>(a), static_cast<T*>(b), kind, wall_t,
                                                         field_t, body_t, g, r);
        break;
    case Variant::EVEN:
        allow_lds(k_aa_even<L, T, FM, MASK>, cap);
        k_aa_even<L, T, FM, MASK><<<grid, block, cap, s>>>(static_cast<T*>(b), kind, field_t, body_t, g, r);
        break;
    case Variant::ODD:
        allow_lds(k_aa_odd<L, T, FM, MASK>, cap);
        k_aa_odd<L, T, FM, MASK><<<grid, block, cap, s>>>(static_cast<T*>(b), kind, wall_t, field_t, body_t, g, r);
        break;
    case Variant::LDS:
        if constexpr (L::D == 2) {
            const size_t tile = sizeof(T) * (size_t)L::Q * (block.x + 2) * (block.y + 2) + (block.x + 2) * (block.y + 2);
            require(tile <= 160 * 1024, "LDS tile does not fit in 160 KB; use a smaller block");
            const size_t lds = occupancy_lds(groups_per_cu, tile);
            allow_lds(k_lds<L, T, FM, MASK>, lds);
            k_lds<L, T, FM, MASK><<<grid, block, lds, s>>>(static_cast<const T*>(a), static_cast<T*>(b), kind,
                                                           wall_t, field_t, body_t, g, r);
        } else {
            require(false, "the LDS-tiled variant is 2D only");
        }
        break;
    }
    check_launch("stream-collide launch");
}

int run_step(const char* name, Variant v, int lattice, int dtype, const void* a, void* b, const uint8_t* kind,
             const void* wall, const void* field, const void* body, int nx, int ny, int nz, double wp,
             double wm, double fx, double fy, double fz, int force_mode, int mask_mode, int bx, int by,
             int groups_per_cu, void* stream)
{
    return guarded(name, [&] {
        const Grid g = make_grid(lattice, nx, ny, nz);
        const dim3 block = block_of(bx, by, step_max_threads_of(lattice, dtype));
        require(groups_per_cu >= 0, "groups_per_cu must be >= 0 (0: no occupancy cap)");
        require(b != nullptr && (v == Variant::EVEN || v == Variant::ODD || a != nullptr), "null population buffer");
        require(mask_mode == 0 || kind != nullptr, "mask_mode 1 needs the kind array");
        require(force_mode >= 0 && force_mode <= 2, "unknown force mode");
        require(force_mode == 0 || mask_mode == 1 || force_mode == 2, "the IBM field needs the kind array");
        require(force_mode != 2 || body != nullptr, "dense force mode needs a body field");
        require(force_mode != 1 || field != nullptr, "field force mode needs a force field");
        const double force[3] = {fx, fy, fz};
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const Relaxation<T> r = make_relaxation<L, T>(wp, wm, force);
            auto go = [&](auto fm, auto mk) {
                launch_step<L, T, decltype(fm)::value, decltype(mk)::value>(v, a, b, kind, wall, field, body, g, r,
                                                                             block, groups_per_cu, as_stream(stream));
            };
            using U = std::integral_constant<ForceMode, ForceMode::UNIFORM>;
            using Fi = std::integral_constant<ForceMode, ForceMode::FIELD>;
            using De = std::integral_constant<ForceMode, ForceMode::DENSE>;
            using M0 = std::integral_constant<bool, false>;
            using M1 = std::integral_constant<bool, true>;
            if (force_mode == 0) {
                if (mask_mode) go(U{}, M1{}); else go(U{}, M0{});
            } else if (force_mode == 1) {
                go(Fi{}, M1{});
            } else {
                if (mask_mode) go(De{}, M1{}); else go(De{}, M0{});
            }
        });
    });
}

// The pull step with in-step link rules and the force layer (k_pull_links).
int run_pull_links(int lattice, int dtype, const void* src, void* dst, const uint8_t* kind, const void* wall,
                   const void* field, const void* body, int nx, int ny, int nz, double wp, double wm, double fx,
                   double fy, double fz, int force_mode, const uint32_t* rule_node_info,
                   const int32_t* rule_node_first, int rule_nodes, const int64_t* rule_slot,
                   const void* rule_coeff, const void* rule_constant, int64_t n_rules, int rule_m,
                   const int64_t* force_slot, const void* force_coeff, const void* force_constant,
                   const double* force_weight, int64_t n_force, int force_m, double* force_out,
                   double* force_partial, int bx, int by, int groups_per_cu, void* stream)
{
    return guarded("lbm_pull_step_links", [&] {
        const Grid g = make_grid(lattice, nx, ny, nz);
        const dim3 block = block_of(bx, by, step_max_threads_of(lattice, dtype));
        require(groups_per_cu >= 0, "groups_per_cu must be >= 0 (0: no occupancy cap)");
        require(src && dst && kind, "null population buffer or kind array");
        require(force_mode >= 0 && force_mode <= 2, "unknown force mode");
        require(force_mode != 2 || body != nullptr, "dense force mode needs a body field");
        require(force_mode != 1 || field != nullptr, "field force mode needs a force field");
        require(n_rules >= 0 && n_rules < (1ll << 31), "rule count out of range");
        require(rule_nodes >= 0 && (n_rules > 0) == (rule_nodes > 0), "rows need rule nodes and nodes rows");
        require(n_rules == 0 || (rule_node_info && rule_node_first && rule_slot && rule_coeff && rule_constant &&
                                 rule_m >= 1 && rule_m <= 128),
                "link rules need the node list, slot, coeff, constant and 1 <= m <= 128");
        require(reinterpret_cast<uintptr_t>(rule_node_info) % 16 == 0, "rule_node_info must be 16-byte aligned");
        require(n_force >= 0, "force row count out of range");
        const int lanes = bx * by;
        require(lanes >= RULE_CHUNK_MAX || rule_nodes == 0, "the rule layer needs at least 64 threads per workgroup");
        const int groups = n_force > 0 ? force_groups(n_force) : 0;
        if (groups) {
            require(force_slot && force_coeff && force_constant && force_weight && force_m >= 1 && force_m <= 128,
                    "force rows need slot, coeff, constant, weight and 1 <= m <= 128");
            require(lanes % 64 == 0, "the force layer needs a whole number of wavefronts per workgroup");
            require(groups == 1 ? force_out != nullptr : force_partial != nullptr,
                    "force rows need out (up to 2048 rows) or partial (above)");
        }
        const double force[3] = {fx, fy, fz};
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const Relaxation<T> r = make_relaxation<L, T>(wp, wm, force);
            const dim3 tiles = grid_of(g, block);
            const int64_t n_tiles = (int64_t)tiles.x * tiles.y * tiles.z;
            const int chunk = rule_chunk_nodes(rule_nodes, (int)std::min<int64_t>(n_tiles, 1 << 20));
            const RuleRows<T> rr{reinterpret_cast<const uint4*>(rule_node_info), rule_node_first, rule_nodes, chunk,
                                 rule_slot, static_cast<const T*>(rule_coeff), static_cast<const T*>(rule_constant),
                                 n_rules, rule_m};
            const ForceRows<T> fw{force_slot, static_cast<const T*>(force_coeff),
                                  static_cast<const T*>(force_constant), force_weight, n_force, force_m, L::D};
            const int64_t work = (rule_nodes ? rule_chunk_count(rule_nodes, chunk) : 0) + groups;
            require(n_tiles + work < (1ll << 31), "grid too large for a one-dimensional launch");
            const dim3 grid((unsigned)(n_tiles + work));
            size_t needed = 0;
            if (rule_nodes) needed = std::max(needed, sizeof(T) * (size_t)chunk * (L::Q - 1));
            if (groups) needed = std::max(needed, sizeof(double) * 3 * FORCE_WAVES);
            const size_t lds = occupancy_lds(groups_per_cu, needed);
            const hipStream_t s = as_stream(stream);
            auto go = [&](auto fm) {
                constexpr ForceMode FM = decltype(fm)::value;
                allow_lds(k_pull_links<L, T, FM>, lds);
                k_pull_links<L, T, FM><<<grid, block, lds, s>>>(
                    static_cast<const T*>(src), static_cast<T*>(dst), kind, static_cast<const T*>(wall),
                    static_cast<const T*>(field), static_cast<const T*>(body), g, r, rr, fw, groups, force_out,
                    force_partial, (int)tiles.x, (int)tiles.y);
            };
            if (force_mode == 0) go(std::integral_constant<ForceMode, ForceMode::UNIFORM>{});
            else if (force_mode == 1) go(std::integral_constant<ForceMode, ForceMode::FIELD>{});
            else go(std::integral_constant<ForceMode, ForceMode::DENSE>{});
            check_launch("pull step with link rules");
        });
    });
}

}  // namespace lbm

using namespace lbm;

namespace {
thread_local std::string g_last_error;
}

void lbm::set_error(const std::string& message) { g_last_error = message; }

extern "C" {

LBM_API const char* lbm_last_error(void) { return g_last_error.c_str(); }

LBM_API const char* lbm_version(void) { return "lbm_hip 0.1.0 (gfx950)"; }

LBM_API int lbm_lattice_tables(int lattice, int* q, int* c, int* opp, double* w)
{
    return guarded("lbm_lattice_tables", [&] {
        dispatch(lattice, LBM_F64, [&](auto Lt, auto) {
            using L = decltype(Lt);
            *q = L::Q;
            for (int i = 0; i < L::Q; ++i) {
                c[i] = L::cx[i];
                c[L::Q + i] = L::cy[i];
                c[2 * L::Q + i] = L::cz[i];
                opp[i] = L::opp[i];
                w[i] = L::w[i];
            }
        });
    });
}

LBM_API int lbm_lattice_tables_device(int lattice, int dtype, int* c, int* opp, void* w, void* stream)
{
    return guarded("lbm_lattice_tables_device", [&] {
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            k_tables<L, T><<<1, 64, 0, as_stream(stream)>>>(c, opp, static_cast<T*>(w));
            check_launch("lattice table kernel");
        });
    });
}

LBM_API int lbm_pull_step(int lattice, int dtype, const void* f_src, void* f_dst, const uint8_t* kind,
                          const void* wall_u, const void* force_field, const void* body_field, int nx, int ny,
                          int nz, double omega_plus, double omega_minus, double fx, double fy, double fz,
                          int force_mode, int mask_mode, int bx, int by, int groups_per_cu, void* stream)
{
    return run_step("lbm_pull_step", Variant::PULL, lattice, dtype, f_src, f_dst, kind, wall_u, force_field,
                    body_field, nx, ny, nz, omega_plus, omega_minus, fx, fy, fz, force_mode, mask_mode, bx, by,
                    groups_per_cu, stream);
}

LBM_API int lbm_pull_step_links(int lattice, int dtype, const void* f_src, void* f_dst, const uint8_t* kind,
                                const void* wall_u, const void* force_field, const void* body_field, int nx, int ny,
                                int nz, double omega_plus, double omega_minus, double fx, double fy, double fz,
                                int force_mode, const uint32_t* rule_node_info, const int32_t* rule_node_first,
                                int rule_nodes, const int64_t* rule_slot,
                                const void* rule_coeff, const void* rule_constant, int64_t n_rules, int rule_m,
                                const int64_t* force_slot, const void* force_coeff, const void* force_constant,
                                const double* force_weight, int64_t n_force, int force_m, double* force_out,
                                double* force_partial, int bx, int by, int groups_per_cu, void* stream)
{
    return run_pull_links(lattice, dtype, f_src, f_dst, kind, wall_u, force_field, body_field, nx, ny, nz,
                          omega_plus, omega_minus, fx, fy, fz, force_mode, rule_node_info, rule_node_first,
                          rule_nodes, rule_slot, rule_coeff,
                          rule_constant, n_rules, rule_m, force_slot, force_coeff, force_constant, force_weight,
                          n_force, force_m, force_out, force_partial, bx, by, groups_per_cu, stream);
}

LBM_API int lbm_aa_even(int lattice, int dtype, void* f, const uint8_t* kind, const void* force_field,
                        const void* body_field, int nx, int ny, int nz, double omega_plus, double omega_minus,
                        double fx, double fy, double fz, int force_mode, int mask_mode, int bx, int by,
                        int groups_per_cu, void* stream)
{
    return run_step("lbm_aa_even", Variant::EVEN, lattice, dtype, nullptr, f, kind, nullptr, force_field,
                    body_field, nx, ny, nz, omega_plus, omega_minus, fx, fy, fz, force_mode, mask_mode, bx, by,
                    groups_per_cu, stream);
}

LBM_API int lbm_aa_odd(int lattice, int dtype, void* f, const uint8_t* kind, const void* wall_u,
                       const void* force_field, const void* body_field, int nx, int ny, int nz, double omega_plus,
                       double omega_minus, double fx, double fy, double fz, int force_mode, int mask_mode, int bx,
                       int by, int groups_per_cu, void* stream)
{
    return run_step("lbm_aa_odd", Variant::ODD, lattice, dtype, nullptr, f, kind, wall_u, force_field, body_field,
                    nx, ny, nz, omega_plus, omega_minus, fx, fy, fz, force_mode, mask_mode, bx, by, groups_per_cu,
                    stream);
}

LBM_API int lbm_lds_step(int lattice, int dtype, const void* f_src, void* f_dst, const uint8_t* kind,
                         const void* wall_u, const void* force_field, const void* body_field, int nx, int ny,
                         int nz, double omega_plus, double omega_minus, double fx, double fy, double fz,
                         int force_mode, int mask_mode, int bx, int by, int groups_per_cu, void* stream)
{
    if (lattice != LBM_D2Q9) {
        set_error("lbm_lds_step: the LDS-tiled variant is 2D only");
        return LBM_ERR_INVALID;
    }
    return run_step("lbm_lds_step", Variant::LDS, lattice, dtype, f_src, f_dst, kind, wall_u, force_field,
                    body_field, nx, ny, nz, omega_plus, omega_minus, fx, fy, fz, force_mode, mask_mode, bx, by,
                    groups_per_cu, stream);
}

LBM_API int lbm_init_equilibrium(int lattice, int dtype, void* f, const void* drho, const void* ux, const void* uy,
                                 const void* uz, int64_t n, int aa, void* stream)
{
    return guarded("lbm_init_equilibrium", [&] {
        require(n > 0 && n < (1ll << 31), "node count out of range");
        require(f && drho && ux && uy, "null pointer");
        require(lattice != LBM_D3Q19 || uz != nullptr, "3D initialisation needs uz");
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const uint32_t nn = (uint32_t)n;
            k_init<L, T><<<(nn + 255) / 256, 256, 0, as_stream(stream)>>>(
                static_cast<T*>(f), static_cast<const T*>(drho), static_cast<const T*>(ux),
                static_cast<const T*>(uy), static_cast<const T*>(uz), nn, aa);
            check_launch("init kernel");
        });
    });
}

LBM_API int lbm_macroscopic(int lattice, int dtype, const void* f, const uint8_t* kind, const void* force_field,
                            int64_t n, int aa, double fx, double fy, double fz, int force_mode, void* rho, void* u,
                            void* stream)
{
    return guarded("lbm_macroscopic", [&] {
        require(n > 0 && n < (1ll << 31), "node count out of range");
        require(force_mode == 0 || (force_field && kind), "the IBM field needs the kind array");
        const double force[3] = {fx, fy, fz};
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const Relaxation<T> r = make_relaxation<L, T>(1.0, 1.0, force);
            const uint32_t nn = (uint32_t)n;
            const dim3 grid((nn + 255) / 256);
            if (force_mode == 0)
                k_macro<L, T, ForceMode::UNIFORM><<<grid, 256, 0, as_stream(stream)>>>(
                    static_cast<const T*>(f), kind, nullptr, nn, aa, r, static_cast<T*>(rho), static_cast<T*>(u));
            else
                k_macro<L, T, ForceMode::FIELD><<<grid, 256, 0, as_stream(stream)>>>(
                    static_cast<const T*>(f), kind, static_cast<const T*>(force_field), nn, aa, r,
                    static_cast<T*>(rho), static_cast<T*>(u));
            check_launch("macroscopic kernel");
        });
    });
}

LBM_API int lbm_health(int lattice, int dtype, const void* f, const uint8_t* kind, int64_t n, int aa, uint32_t* out,
                       void* stream)
{
    return guarded("lbm_health", [&] {
        require(n > 0 && n < (1ll << 31), "node count out of range");
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const uint32_t nn = (uint32_t)n;
            k_health<L, T><<<(nn + 255) / 256, 256, 0, as_stream(stream)>>>(static_cast<const T*>(f), kind, nn, aa,
                                                                            out);
            check_launch("health kernel");
        });
    });
}

LBM_API int lbm_kernel_resources(int lattice, int dtype, int* values, int max_kernels, char* names, int names_len,
                                 int* count)
{
    return guarded("lbm_kernel_resources", [&] {
        std::string joined;
        int k = 0;
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            auto add = [&](const char* name, const void* fn) {
                hipFuncAttributes attr{};
                hip_check(hipFuncGetAttributes(&attr, fn), "hipFuncGetAttributes");
                if (k < max_kernels) {
                    values[4 * k + 0] = attr.numRegs;
                    values[4 * k + 1] = (int)attr.localSizeBytes;
                    values[4 * k + 2] = (int)attr.sharedSizeBytes;
                    values[4 * k + 3] = attr.maxThreadsPerBlock;
                }
                joined += (k ? ";" : "");
                joined += name;
                ++k;
            };
            add("pull", reinterpret_cast<const void*>(&k_pull<L, T, ForceMode::UNIFORM, true>));
            add("pull_all_fluid", reinterpret_cast<const void*>(&k_pull<L, T, ForceMode::UNIFORM, false>));
            add("pull_links", reinterpret_cast<const void*>(&k_pull_links<L, T, ForceMode::UNIFORM>));
            add("pull_links_field", reinterpret_cast<const void*>(&k_pull_links<L, T, ForceMode::FIELD>));
            add("aa_even", reinterpret_cast<const void*>(&k_aa_even<L, T, ForceMode::UNIFORM, true>));
            add("aa_odd", reinterpret_cast<const void*>(&k_aa_odd<L, T, ForceMode::UNIFORM, true>));
            add("aa_even_all_fluid", reinterpret_cast<const void*>(&k_aa_even<L, T, ForceMode::UNIFORM, false>));
            add("aa_odd_all_fluid", reinterpret_cast<const void*>(&k_aa_odd<L, T, ForceMode::UNIFORM, false>));
            add("aa_odd_field", reinterpret_cast<const void*>(&k_aa_odd<L, T, ForceMode::FIELD, true>));
            add("aa_odd_dense", reinterpret_cast<const void*>(&k_aa_odd<L, T, ForceMode::DENSE, true>));
            if constexpr (L::D == 2) add("lds", reinterpret_cast<const void*>(&k_lds<L, T, ForceMode::UNIFORM, true>));
        });
        *count = k;
        if (names && names_len > 0) {
            std::strncpy(names, joined.c_str(), (size_t)names_len - 1);
            names[names_len - 1] = '\0';
        }
    });
}

}  // extern "C"
