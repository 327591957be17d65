// The link-rule layer of the pull step: rule rows evaluated inside the step by the workgroups
// that update the rule nodes (used by the pull step with link rules and by the fused adjoint
// step's forward recompute).
#pragma once

#include "lbm_affine.h"
#include "lbm_common.h"
#include "lbm_step.h"

namespace lbm {

// The link rules of the two-lattice step, evaluated inside the step. A rule of link (x, j)
// (x - c_j solid) gives f_j(x, t + 1) as an affine row of the stored state; the pre-pass form
// writes it into the source slot (opp j, x) that the bounce-back of direction j reads. The
// pull step never writes its source buffer, so rows evaluated during the step read exactly
// the pre-rule populations the pre-pass reads, and the step can use the values without the
// pre-pass launch. Rows are sorted by (node, direction); the nodes with rows (sorted) are
// described by node_info[k] = {node index, directions i whose upstream node x - c_i is
// solid, directions j with a row, the node's first row}, and node_first[k] = that first
// row again (node_first[count] = n) gives a chunk's row range. The nodes are updated in
// chunks of `chunk` nodes (rule_chunk_nodes).
template <class T>
struct RuleRows {
    const uint4* node_info;
    const int32_t* node_first;
    int count;
    int chunk;
    const int64_t* slot;
    const T* coeff;
    const T* constant;
    int64_t n;
    int m;
};

// Rule nodes per chunk: at most one wavefront of node lanes, and few enough that the chunks
// cover the layer's workgroups once where they can (at least 16 nodes each), so the nodes'
// scattered loads and stores spread over many compute units instead of queueing in a few
// (880 x 164 Bouzidi channel, 492 rule nodes on 164 workgroups: 16-node chunks). Chunks
// also grow until there are at most RULE_CHUNKS_MAX of them: the chunk workgroups come first
// in the grid, and thousands of them, each waiting on its dependent loads, hold back the
// tiles (D3Q19 duct 384 x 96 x 96, 19,154 rule nodes: 146.7 us per pull step with 16-node
// chunks, 141.4 us with 64).
constexpr int RULE_CHUNK_MIN = 16;
constexpr int RULE_CHUNK_MAX = 64;
constexpr int RULE_CHUNKS_MAX = 512;

inline int rule_chunk_nodes(int count, int hosts)
{
    int chunk = RULE_CHUNK_MIN;
    while (chunk < RULE_CHUNK_MAX &&
           ((int64_t)chunk * hosts < count || (int64_t)chunk * RULE_CHUNKS_MAX < count))
        chunk *= 2;
    return chunk;
}

__host__ __device__ constexpr int rule_chunk_count(int count, int chunk) { return (count + chunk - 1) / chunk; }


// One chunk of rule nodes by one workgroup of `lanes` threads (tid the linear index): the
// workgroup evaluates the chunk's rows into the LDS array `vals` (row first + e by thread
// e mod lanes), threads below rr.chunk update node c rr.chunk + tid as pull_update does,
// their rule directions taking the rows' values after the barrier. With the blocked set in
// the node's record, each population is loaded once from its selected slot (pulled or
// bounce-back), in flight with the rows' loads, all before the barrier. Same operations in
// the same order as pull_update after a pre-pass: the same bits. WRITE_BACK also stores each
// row's value into the source slot the pre-pass would have written, so the source buffer ends
// as the pre-pass leaves it (for rule sets in which no row reads another row's target).
template <class L, class T, ForceMode FM, bool WRITE_BACK = false>
__device__ __forceinline__ void rule_chunk(const T* __restrict__ src, T* __restrict__ dst,
                                           const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                           const T* __restrict__ field, const T* __restrict__ body, const Grid& g,
                                           const Relaxation<T>& r, const RuleRows<T>& rr, int c, int lanes, int tid,
                                           T* __restrict__ vals)
{
    const int first_node = c * rr.chunk;
    const int end_node = min(first_node + rr.chunk, rr.count);
    const bool node_lane = first_node + tid < end_node;
    uint4 info = make_uint4(0u, 0u, 0u, 0u);
    if (node_lane) info = rr.node_info[first_node + tid];
    const int64_t first_row = rr.node_first[first_node];
    const int64_t end_row = rr.node_first[end_node];
    Coord cc{0, 0, 0};
    const uint32_t node = info.x;
    const uint32_t blocked = info.y;
    uint8_t k = 0;
    T f[L::Q];
    static_for<0, L::Q>([&](auto I) { f[I] = T(0); });
    if (node_lane) {
        cc.x = (int)(node % (uint32_t)g.nx);
        const uint32_t rest = node / (uint32_t)g.nx;
        cc.y = (int)(rest % (uint32_t)g.ny);
        cc.z = (int)(rest / (uint32_t)g.ny);
        k = kind[node];
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            const bool bounce = (blocked >> i) & 1u;
            f[i] = src[bounce ? (size_t)L::opp[i] * g.ns + node
                              : (size_t)i * g.ns + neighbour<L, i, -1>(cc.x, cc.y, cc.z, g)];
        });
    }
    for (int64_t e = tid; first_row + e < end_row; e += lanes)
        vals[e] = affine_row(src, rr.slot, rr.coeff, rr.constant, rr.n, rr.m, first_row + e);
    __syncthreads();
    if (node_lane) {
        uint32_t has = info.z;
        int local = (int)((int64_t)info.w - first_row);
        while (has) {
            const int j = __ffs(has) - 1;
            has &= has - 1u;
            const T v = vals[local++];
            // select into every slot (a store at a runtime index would move f to scratch)
            static_for<1, L::Q>([&](auto I) { f[I] = (j == int(I)) ? v : f[I]; });
            // The pre-pass form's state: the value in the source slot (opp j, node) it would
            // have written. Only this node reads that slot (its upstream node is solid), and
            // it did so before the barrier; no row reads another row's target.
            if constexpr (WRITE_BACK) const_cast<T*>(src)[(size_t)L::opp[j] * g.ns + node] = v;
        }
        add_wall_terms<L, -1>(f, blocked, wall_u, cc, g);
        T Fx = T(0), Fy = T(0), Fz = T(0);
        if constexpr (FM != ForceMode::UNIFORM) node_force<L, FM>(r, k, field, body, node, g.n, Fx, Fy, Fz);
        collide<L, FM>(f, r, Fx, Fy, Fz);
        static_for<0, L::Q>([&](auto I) { dst[(size_t)I * g.ns + node] = f[I]; });
    }
    __syncthreads();  // vals is reused by the next chunk
}

}  // namespace lbm
