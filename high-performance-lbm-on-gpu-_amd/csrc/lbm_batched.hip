// Many independent lattices of uneven sizes advanced by one persistent launch per step.
//
// No counterpart in the reference (one lattice per solver there). All lattices keep their
// populations in one pool (lattice l at element offset f_offset[l], its own SoA planes),
// their kind bytes and wall velocities in two more pools. A work table lists every
// (lattice, tile), a tile being a run of consecutive node indices of one lattice as long
// as the workgroup (rows wrap inside a tile, so lattices of any width leave no lane idle
// except in their last tile); a fixed grid of workgroups walks the table with a static
// stride (workgroup w takes tiles w, w + G, w + 2G, ...), so the assignment is
// deterministic and no atomics are involved. Each node runs pull_update, the code of the
// single-lattice pull kernel, with its lattice's own grid and relaxation constants, so a
// batched step produces the same bits as running every lattice alone.

#include "lbm_common.h"
#include "lbm_step.h"

namespace lbm {

template <class T>
struct LatticeDesc {
    Grid g;
    int pad;
    uint64_t f_offset;     // elements into the population pools (q * g.ns per lattice)
    uint64_t kind_offset;  // bytes into the kind pool (n per lattice)
    uint64_t wall_offset;  // elements into the wall pool (d * n per lattice)
    Relaxation<T> r;
};

// Population planes in the pool start every POOL_PLANE_ALIGN elements (each lattice's
// node count rounded up), so every plane of every lattice begins on a cache-line boundary
// and a wavefront's unshifted access covers two 128-byte lines, not three.
constexpr uint32_t POOL_PLANE_ALIGN = 64;

struct TileDesc {
    int lattice;
    uint32_t start;  // first node index (within the lattice) of the tile
};

template <class L, class T, bool MASK>
__global__ __launch_bounds__(step_max_threads(L::D, sizeof(T))) void k_batched_pull(const T* __restrict__ src, T* __restrict__ dst,
                                                       const uint8_t* __restrict__ kind, const T* __restrict__ wall,
                                                       const LatticeDesc<T>* __restrict__ lattices,
                                                       const TileDesc* __restrict__ tiles, int n_tiles)
{
    const uint32_t lane = threadIdx.y * blockDim.x + threadIdx.x;
    int t = blockIdx.x;
    if (t >= n_tiles) return;
    TileDesc td = tiles[t];
    for (;;) {
        // The next tile's descriptor is requested (from an index valid either way) before
        // this tile's update, so the table read overlaps this tile's population traffic
        // instead of stalling every wavefront of the workgroup between tiles.
        const int tn = t + (int)gridDim.x;
        const bool more = tn < n_tiles;
        const TileDesc next = tiles[more ? tn : t];
        // Tile and lattice descriptors are the same for the whole workgroup: kept in scalar
        // registers, the lattice descriptor is one scalar read, not per-lane vector loads.
        const int lat = __builtin_amdgcn_readfirstlane(td.lattice);
        const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane((int)td.start);
        const LatticeDesc<T>& d = lattices[lat];
        const Grid g = d.g;
        const uint32_t node = first + lane;
        if (node < g.n) {
            const uint32_t row = node / (uint32_t)g.nx;
            const Coord c{(int)(node - row * (uint32_t)g.nx), (int)(row % (uint32_t)g.ny), (int)(row / (uint32_t)g.ny)};
            pull_update<L, T, ForceMode::UNIFORM, MASK>(src + d.f_offset, dst + d.f_offset, kind + d.kind_offset,
                                                       wall ? wall + d.wall_offset : nullptr, nullptr, nullptr, g, d.r, c);
        }
        if (!more) break;
        t = tn;
        td = next;
    }
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LBM_API int lbm_batched_desc_size(int dtype, int64_t* bytes)
{
    return guarded("lbm_batched_desc_size", [&] {
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            *bytes = (int64_t)sizeof(LatticeDesc<T>);
        });
    });
}

LBM_API int lbm_batched_describe(int lattice, int dtype, int n_lattices, const int* nx, const int* ny,
                                 const int* nz, const double* omega_plus, const double* omega_minus,
                                 const double* force, void* host_desc)
{
    return guarded("lbm_batched_describe", [&] {
        require(n_lattices > 0 && nx && ny && nz && omega_plus && omega_minus && force && host_desc, "bad arguments");
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            auto* out = static_cast<LatticeDesc<T>*>(host_desc);
            uint64_t f = 0, k = 0, w = 0;
            for (int l = 0; l < n_lattices; ++l) {
                LatticeDesc<T> d{};
                d.g = make_grid(lattice, nx[l], ny[l], nz[l]);
                d.g.ns = (uint32_t)(((uint64_t)d.g.n + POOL_PLANE_ALIGN - 1) / POOL_PLANE_ALIGN * POOL_PLANE_ALIGN);
                d.f_offset = f;
                d.kind_offset = k;
                d.wall_offset = w;
                d.r = make_relaxation<L, T>(omega_plus[l], omega_minus[l], force + 3 * l);
                f += (uint64_t)L::Q * d.g.ns;
                k += d.g.n;
                w += (uint64_t)L::D * d.g.n;
                out[l] = d;
            }
        });
    });
}

LBM_API int lbm_batched_pull_step(int lattice, int dtype, const void* src_pool, void* dst_pool,
                                  const uint8_t* kind_pool, const void* wall_pool, const void* desc,
                                  const int* tiles, int n_tiles, int mask_mode, int bx, int by, int groups,
                                  int groups_per_cu, void* stream)
{
    return guarded("lbm_batched_pull_step", [&] {
        require(n_tiles > 0 && src_pool && dst_pool && desc && tiles, "bad arguments");
        require(mask_mode == 0 || kind_pool, "mask_mode 1 needs the kind pool");
        require(groups > 0, "groups must be positive");
        require(groups_per_cu >= 0, "groups_per_cu must be >= 0 (0: no occupancy cap)");
        const size_t cap = occupancy_lds(groups_per_cu, 0);
        const dim3 block = block_of(bx, by, step_max_threads_of(lattice, dtype));
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const unsigned grid = (unsigned)std::min(groups, n_tiles);
            auto* td = reinterpret_cast<const TileDesc*>(tiles);
            auto* ld = static_cast<const LatticeDesc<T>*>(desc);
            auto go = [&](auto mk) {
                constexpr bool MK = decltype(mk)::value;
                allow_lds(k_batched_pull<L, T, MK>, cap);
                k_batched_pull<L, T, MK><<<grid, block, cap, as_stream(stream)>>>(
                    static_cast<const T*>(src_pool), static_cast<T*>(dst_pool), kind_pool,
                    static_cast<const T*>(wall_pool), ld, td, n_tiles);
            };
            if (mask_mode) go(std::true_type{});
            else go(std::false_type{});
            check_launch("batched pull step");
        });
    });
}

}  // extern "C"
