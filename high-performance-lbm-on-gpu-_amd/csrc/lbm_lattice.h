// Lattice constants as constexpr tables, plus a compile-time loop helper.
//
// D2Q9 and D3Q19 in the reference's direction numbering (src/lbm/lattice.py:114-146).
// Every loop over directions in the kernels is a static_for, so the direction
// index is a compile-time constant: the tables fold into immediates and the Q
// populations of a node live in VGPRs (no scratch).
#pragma once

#include <cstdint>
#include <type_traits>
#include <utility>

namespace lbm {

template <int N>
using ic = std::integral_constant<int, N>;

template <int Begin, int End, class F>
__host__ __device__ __forceinline__ constexpr void static_for(F&& f)
{
    if constexpr (Begin < End) {
        f(ic<Begin>{});
        static_for<Begin + 1, End>(static_cast<F&&>(f));
    }
}

struct D2Q9 {
    static constexpr int D = 2;
    static constexpr int Q = 9;
    static constexpr int cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    static constexpr double w[Q] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int code = 0;
};

struct D3Q19 {
    static constexpr int D = 3;
    static constexpr int Q = 19;
    static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    static constexpr double w[Q] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                    1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int code = 1;
};

template <class L, int I>
struct Dir {
    static constexpr int x = L::cx[I];
    static constexpr int y = L::cy[I];
    static constexpr int z = L::cz[I];
    static constexpr int o = L::opp[I];
    static constexpr double w = L::w[I];
    // 2 w / cs^2 with cs^2 = 1/3, evaluated in double exactly as the host does.
    static constexpr double wall = 2.0 * L::w[I] / (1.0 / 3.0);
    static constexpr bool rest = (x == 0 && y == 0 && z == 0);
};

// Node kind bits (one byte per node). SOLID and FORCED have the reference's
// values (0 fluid, 1 solid, 2 fluid with a local IBM force); NEAR_SOLID is set
// on the host for fluid nodes with a solid node at some x - c_i, so interior
// wavefronts skip the per-direction solid checks.
constexpr uint8_t KIND_SOLID = 1;
constexpr uint8_t KIND_FORCED = 2;
constexpr uint8_t KIND_NEAR_SOLID = 4;
// A fluid node with link rules that lbm_pull_step_links updates in its rule layer (the
// host sets it on exactly the nodes of the rule-node list it passes; other kernels ignore it).
constexpr uint8_t KIND_RULES = 8;
// A node that a fused launch updates in a workgroup of its own, so the tile workgroups skip
// it: in the fused backward step the IBM stencil nodes and the link-rule nodes, in the fused
// forward step (k_ibm_step) the IBM stencil nodes. Set only in the fused steps' own copies
// of the kind bytes.
constexpr uint8_t KIND_ADJ_REGION = 16;

}  // namespace lbm
