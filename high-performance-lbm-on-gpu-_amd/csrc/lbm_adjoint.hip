// Discrete adjoint of the pull time step: transposed collision Jacobian + push of the
// adjoint populations, transposed link rules, and the adjoint of the fused IBM forcing.
//
// Forward step (pull layout, the state buffer already carries the link-rule values):
//   f_s(x)   = f^t(x - c_i)  or  f^t_opp(i)(x) + wall term   (bounce-back)
//   f^{t+1}  = C(f_s(x), F(x)),  F = F0 + g(x) + h(x)          (h: IBM field at forced nodes)
// Backward step for lam = dL/df^{t+1}:
//   mu   = (dC/df_s)^T lam(x),  nu = (dC/dF)^T lam(x)          (per node, reverse mode)
//   lam^t(x - c_i)[i] = mu_i(x), or lam^t(x)[opp i] = mu_i(x) when x - c_i is solid
// Every slot (k, y) of a fluid node y is written exactly once per step (from the node
// y + c_k, or from y itself when y + c_k is solid), the same exclusivity argument as the
// AA odd step, so the push needs no atomics; solid-node slots are never written and
// stay zero. dL/dg accumulates nu in float64.

#include "lbm_affine.h"
#include "lbm_common.h"
#include "lbm_ibm_core.h"
#include "lbm_rule_layer.h"
#include "lbm_step.h"

#include <initializer_list>
#include <type_traits>

namespace lbm {

// Reverse mode of collide<L, FM>: given f (streamed), F and lam = dL/df*, return
// fbar = dL/df and (Fbx, Fby, Fbz) = dL/dF. Mirrors the forward expressions:
//   f*_i = f_i - (hp + hm) n_i - (hp - hm) n_opp + S_i,  n = f - feq,
//   feq_i = w_i (drho + rho g_i),  g_i = 3 cu_i + 4.5 cu_i^2 - 1.5 u.u,
//   S_i = w_i [a+ (9 cu_i cF_i - 3 u.F) + a- 3 cF_i],  u = (j + F/2) / rho.
template <class L, ForceMode FM, class T>
__device__ __forceinline__ void collide_adjoint(const T (&f)[L::Q], const T (&lam)[L::Q], const Relaxation<T>& r,
                                                T Fx, T Fy, T Fz, T (&fbar)[L::Q], T& Fbx, T& Fby, T& Fbz)
{
    T drho, jx, jy, jz;
    moments<L>(f, drho, jx, jy, jz);
    const T rho = T(1) + drho;
    T Gx, Gy, Gz;
    if constexpr (FM == ForceMode::UNIFORM) {
        Gx = r.force[0];
        Gy = r.force[1];
        Gz = r.force[2];
    } else {
        Gx = Fx;
        Gy = Fy;
        Gz = Fz;
    }
    // One reciprocal for the seven quotients by rho (the forward divides; the adjoint's
    // values agree with it to rounding, which the gradient checks allow for).
    const T inv_rho = T(1) / rho;
    const T ux = (jx + T(0.5) * Gx) * inv_rho;
    const T uy = (jy + T(0.5) * Gy) * inv_rho;
    const T uz = L::D == 3 ? (jz + T(0.5) * Gz) * inv_rho : T(0);
    const T alpha = r.half_wp + r.half_wm;
    const T beta = r.half_wp - r.half_wm;
    const bool guo = (FM != ForceMode::UNIFORM) || r.has_force;

    T drho_bar = T(0), rho_bar = T(0), gsum = T(0);
    T ubx = T(0), uby = T(0), ubz = T(0);
    T fbx = T(0), fby = T(0), fbz = T(0);
    const T usq = T(1.5) * fma(ux, ux, fma(uy, uy, uz * uz));
    // c_i . v from the compile-time components (no products by 0 or +-1).
    auto dot = [](auto I, T vx, T vy, T vz) {
        using Di = Dir<L, decltype(I)::value>;
        T d = T(0);
        bool first = true;
        auto add = [&](int c, T v) {
            if (c == 0) return;
            const T t = c > 0 ? v : -v;
            d = first ? t : d + t;
            first = false;
        };
        add(Di::x, vx);
        add(Di::y, vy);
        if constexpr (L::D == 3) add(Di::z, vz);
        return d;
    };
    // a += c_i,axis * v for the three axes (skipped where the component is 0).
    auto push = [](auto I, T& ax, T& ay, T& az, T v) {
        using Di = Dir<L, decltype(I)::value>;
        if constexpr (Di::x != 0) ax = Di::x > 0 ? ax + v : ax - v;
        if constexpr (Di::y != 0) ay = Di::y > 0 ? ay + v : ay - v;
        if constexpr (L::D == 3 && Di::z != 0) az = Di::z > 0 ? az + v : az - v;
    };
    // The expressions below are those of the comment above with products and sums fused
    // (the adjoint is checked against float64 finite differences and autograd, not bit
    // for bit against a host code).
    const T G3x = T(3) * Gx, G3y = T(3) * Gy, G3z = T(3) * Gz;
    const T am3 = T(3) * r.a_minus;
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        using Di = Dir<L, i>;
        const T nbar = fma(-alpha, lam[i], -(beta * lam[L::opp[i]]));
        fbar[i] = lam[i] + nbar;
        const T fw = -nbar * T(Di::w);  // feqbar w
        const T cu = dot(I, ux, uy, uz);
        const T g = fma(T(4.5) * cu, cu, fma(T(3), cu, -usq));
        drho_bar += fw;
        rho_bar = fma(fw, g, rho_bar);
        const T gbar = fw * rho;
        gsum += gbar;
        push(I, ubx, uby, ubz, gbar * fma(T(9), cu, T(3)));
        if (guo) {
            const T cF = dot(I, Gx, Gy, Gz);
            const T lw = lam[i] * T(Di::w);
            const T lp = lw * r.a_plus;
            ubx = fma(-lp, G3x, ubx);
            uby = fma(-lp, G3y, uby);
            if constexpr (L::D == 3) ubz = fma(-lp, G3z, ubz);
            push(I, ubx, uby, ubz, T(9) * cF * lp);
            fbx = fma(-lp, T(3) * ux, fbx);
            fby = fma(-lp, T(3) * uy, fby);
            if constexpr (L::D == 3) fbz = fma(-lp, T(3) * uz, fbz);
            push(I, fbx, fby, fbz, fma(T(9) * cu, lp, lw * am3));
        }
    });
    const T g3 = T(3) * gsum;
    ubx = fma(-g3, ux, ubx);
    uby = fma(-g3, uy, uby);
    ubz = L::D == 3 ? fma(-g3, uz, ubz) : T(0);
    const T jbx = ubx * inv_rho, jby = uby * inv_rho, jbz = ubz * inv_rho;
    fbx = fma(T(0.5), jbx, fbx);
    fby = fma(T(0.5), jby, fby);
    fbz = fma(T(0.5), jbz, fbz);
    rho_bar = fma(-fma(ubx, ux, fma(uby, uy, ubz * uz)), inv_rho, rho_bar);
    drho_bar += rho_bar;
    static_for<0, L::Q>([&](auto I) { fbar[I] += drho_bar + dot(I, jbx, jby, jbz); });
    Fbx = fbx;
    Fby = fby;
    Fbz = L::D == 3 ? fbz : T(0);
}

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

template <class L, int I, int S>
__device__ __forceinline__ uint32_t nbr(int x, int y, int z, const Grid& g)
{
    using Di = Dir<L, I>;
    const int xx = Di::x == 0 ? x : wrapi(x + S * Di::x, g.nx);
    const int yy = Di::y == 0 ? y : wrapi(y + S * Di::y, g.ny);
    const int zz = (L::D == 2 || Di::z == 0) ? z : wrapi(z + S * Di::z, g.nz);
    return ((uint32_t)zz * (uint32_t)g.ny + (uint32_t)yy) * (uint32_t)g.nx + (uint32_t)xx;
}

// The adjoint update of node (x, y, z) (inside the grid). SKIP_REGION leaves the nodes marked
// KIND_ADJ_REGION to the fused step's region workgroup.
template <class L, class T, ForceMode FM, bool MASK, bool SKIP_REGION = false>
__device__ __forceinline__ void adj_pull_node(const T* __restrict__ lam_next, T* __restrict__ lam_out,
                                              const T* __restrict__ f, const uint8_t* __restrict__ kind,
                                              const T* __restrict__ wall_u, const T* __restrict__ field,
                                              const T* __restrict__ body, double* __restrict__ grad_body,
                                              T* __restrict__ nu_forced, const Grid& g, const Relaxation<T>& r, int x,
                                              int y, int z)
{
    const uint32_t node = ((uint32_t)z * (uint32_t)g.ny + (uint32_t)y) * (uint32_t)g.nx + (uint32_t)x;
    // The kind byte, the Q pulled populations and the Q adjoint populations are loaded
    // together (valid addresses for any node; a solid node discards them). A node with a
    // solid upstream neighbour then reloads each population from its selected slot
    // (bounce-back or pulled) and adds the moving-wall terms; stores below also go to
    // selected slots, so all loads and stores of a pass stay in flight together.
    const uint8_t k = MASK ? kind[node] : 0;
    T fs[L::Q], lam[L::Q];
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        fs[i] = f[(size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)];
        lam[i] = lam_next[(size_t)i * g.n + node];
    });
    if constexpr (MASK) {
        hold_loaded(fs);
        hold_loaded(lam);
    }
    if (MASK && (k & (SKIP_REGION ? (KIND_SOLID | KIND_ADJ_REGION) : KIND_SOLID))) return;
    uint32_t blocked = 0;
    if (MASK && (k & KIND_NEAR_SOLID)) {
        uint32_t nk[L::Q];  // all neighbour kinds read together, held, then tested
        nk[0] = 0;
        static_for<1, L::Q>([&](auto I) { nk[I] = kind[nbr<L, I, -1>(x, y, z, g)]; });
        hold_loaded(nk);
        static_for<1, L::Q>([&](auto I) {
            if (nk[I] & KIND_SOLID) blocked |= 1u << I;
        });
    }
    if (MASK && blocked) {
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            const bool bounce = (blocked >> i) & 1u;
            fs[i] = f[bounce ? (size_t)L::opp[i] * g.n + node : (size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)];
        });
        add_wall_terms<L, -1>(fs, blocked, wall_u, Coord{x, y, z}, g);
    }
    T Fx = r.force[0], Fy = r.force[1], Fz = r.force[2];
    if constexpr (FM == ForceMode::DENSE) {
        Fx = Fx + body[node];
        Fy = Fy + body[(size_t)g.n + node];
        if constexpr (L::D == 3) Fz = Fz + body[2 * (size_t)g.n + node];
    }
    if constexpr (FM != ForceMode::UNIFORM) {
        if (k & KIND_FORCED) {
            Fx = Fx + field[node];
            Fy = Fy + field[(size_t)g.n + node];
            if constexpr (L::D == 3) Fz = Fz + field[2 * (size_t)g.n + node];
        }
    }
    T mu[L::Q];
    T Fbx, Fby, Fbz;
    collide_adjoint<L, FM>(fs, lam, r, Fx, Fy, Fz, mu, Fbx, Fby, Fbz);
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        const bool bounce = MASK && ((blocked >> i) & 1u);
        lam_out[bounce ? (size_t)L::opp[i] * g.n + node : (size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)] = mu[i];
    });
    if (grad_body) {
        grad_body[node] += (double)Fbx;
        grad_body[(size_t)g.n + node] += (double)Fby;
        if constexpr (L::D == 3) grad_body[2 * (size_t)g.n + node] += (double)Fbz;
    }
    if (nu_forced && (k & KIND_FORCED)) {
        nu_forced[node] = Fbx;
        nu_forced[(size_t)g.n + node] = Fby;
        if constexpr (L::D == 3) nu_forced[2 * (size_t)g.n + node] = Fbz;
    }
}

// The adjoint update of one fluid node (adj_pull_node) in two parts, so that all of its
// loads go out in one level, with other loads: load() reads the kind byte, the pulled and the
// adjoint populations, the node's forces and its float64 body-force gradients (no other
// workgroup writes them in the launch); finish() is the rest of adj_pull_node (same
// operations, so the same bits). Without MASKED the node must have no solid upstream
// neighbour; with MASKED load() takes the set of directions whose upstream node is solid
// (bit i for x - c_i), loads those populations from their bounce-back slots and the wall
// terms, and the pushes of those directions go to the bounce-back slots.
template <class L, class T, ForceMode FM, bool MASKED = false>
struct AdjNodeParts {
    static constexpr int QM = MASKED ? L::Q : 1;
    uint32_t node, bits;
    int x, y, z;
    uint8_t k;
    T fs[L::Q], lam[L::Q];
    T body[3], field[3];
    double gb[3];
    T wt[QM];

    __device__ __forceinline__ void load(const T* __restrict__ lam_next, const T* __restrict__ f,
                                         const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                         const T* __restrict__ fld, const T* __restrict__ bdy,
                                         const double* __restrict__ grad_body, const Grid& g, uint32_t at,
                                         uint32_t blocked = 0)
    {
        node = at;
        bits = MASKED ? blocked : 0u;
        const uint32_t rest = node / (uint32_t)g.nx;
        x = (int)(node - rest * (uint32_t)g.nx);
        y = (int)(rest % (uint32_t)g.ny);
        z = (int)(rest / (uint32_t)g.ny);
        k = kind[node];
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            const bool bounce = MASKED && ((bits >> i) & 1u);
            fs[i] = f[bounce ? (size_t)L::opp[i] * g.n + node : (size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)];
            lam[i] = lam_next[(size_t)i * g.n + node];
        });
#pragma unroll
        for (int a = 0; a < L::D; ++a) {
            body[a] = FM == ForceMode::DENSE ? bdy[(size_t)a * g.n + node] : T(0);
            field[a] = FM != ForceMode::UNIFORM ? fld[(size_t)a * g.n + node] : T(0);
            gb[a] = grad_body ? grad_body[(size_t)a * g.n + node] : 0.0;
        }
        if constexpr (MASKED) {
            if (wall_u) {
                static_for<0, L::Q>([&](auto I) { wt[I] = wall_term<L, I>(wall_u, neighbour<L, I, -1>(x, y, z, g), g.n); });
            }
        }
    }

    // mu (the adjoint pushes) and dL/dF of the node; blocked directions take the wall terms.
    __device__ __forceinline__ void adjoint(const T* __restrict__ wall_u, const Relaxation<T>& r, T (&mu)[L::Q],
                                            T& Fbx, T& Fby, T& Fbz)
    {
        if constexpr (MASKED) {
            if (wall_u) {
                static_for<0, L::Q>([&](auto I) {
                    if ((bits >> I) & 1u) fs[I] = fs[I] + wt[I];
                });
            }
        }
        T Fx = r.force[0], Fy = r.force[1], Fz = r.force[2];
        if constexpr (FM == ForceMode::DENSE) {
            Fx = Fx + body[0];
            Fy = Fy + body[1];
            if constexpr (L::D == 3) Fz = Fz + body[2];
        }
        if constexpr (FM != ForceMode::UNIFORM) {
            if (k & KIND_FORCED) {
                Fx = Fx + field[0];
                Fy = Fy + field[1];
                if constexpr (L::D == 3) Fz = Fz + field[2];
            }
        }
        collide_adjoint<L, FM>(fs, lam, r, Fx, Fy, Fz, mu, Fbx, Fby, Fbz);
    }

    __device__ __forceinline__ void gradients(double* __restrict__ grad_body, const Grid& g, T Fbx, T Fby, T Fbz) const
    {
        if (grad_body) {
            grad_body[node] = gb[0] + (double)Fbx;
            grad_body[(size_t)g.n + node] = gb[1] + (double)Fby;
            if constexpr (L::D == 3) grad_body[2 * (size_t)g.n + node] = gb[2] + (double)Fbz;
        }
    }

    // The pushes mu to their slots, the body-force gradients and nu at a forced node.
    __device__ __forceinline__ void store(T* __restrict__ lam_out, const T (&mu)[L::Q], double* __restrict__ grad_body,
                                          T* __restrict__ nu_forced, const Grid& g, T Fbx, T Fby, T Fbz) const
    {
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            const bool bounce = MASKED && ((bits >> i) & 1u);
            lam_out[bounce ? (size_t)L::opp[i] * g.n + node : (size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)] = mu[i];
        });
        gradients(grad_body, g, Fbx, Fby, Fbz);
        if (nu_forced && (k & KIND_FORCED)) {
            nu_forced[node] = Fbx;
            nu_forced[(size_t)g.n + node] = Fby;
            if constexpr (L::D == 3) nu_forced[2 * (size_t)g.n + node] = Fbz;
        }
    }

    __device__ __forceinline__ void finish(T* __restrict__ lam_out, double* __restrict__ grad_body,
                                           T* __restrict__ nu_forced, const T* __restrict__ wall_u, const Grid& g,
                                           const Relaxation<T>& r)
    {
        T mu[L::Q];
        T Fbx, Fby, Fbz;
        adjoint(wall_u, r, mu, Fbx, Fby, Fbz);
        store(lam_out, mu, grad_body, nu_forced, g, Fbx, Fby, Fbz);
    }

    // finish() for an IBM node that keeps its pushes: mu_i stays in lam[i] (the node adds the
    // transposed moment rows itself in push_kept) and dL/dF (the node being forced) goes to
    // nu_seed instead of nu; the body-force gradients are summed as in finish() and stored
    // by push_kept (a store here would hold the next barrier until it completes).
    __device__ __forceinline__ void finish_kept(T (&nu_seed)[L::D], const Relaxation<T>& r)
    {
        static_assert(!MASKED, "an IBM node has no solid upstream neighbour");
        T mu[L::Q];
        T Fbx, Fby, Fbz;
        adjoint(nullptr, r, mu, Fbx, Fby, Fbz);
        static_for<0, L::Q>([&](auto I) { lam[I] = mu[I]; });
        gb[0] = gb[0] + (double)Fbx;
        gb[1] = gb[1] + (double)Fby;
        if constexpr (L::D == 3) gb[2] = gb[2] + (double)Fbz;
        const bool forced = (k & KIND_FORCED) != 0;
        nu_seed[0] = forced ? Fbx : T(0);
        nu_seed[1] = forced ? Fby : T(0);
        if constexpr (L::D == 3) nu_seed[2] = forced ? Fbz : T(0);
    }

    // The kept pushes plus the transposed moment rows of IBM node `row` (rows a * nn + row,
    // a = 0 density with coefficient 1, a = 1 .. d momentum with c_i,a-1, the nonzero ones in
    // increasing a, summed from 0 as the transposed row lists are): push i writes slot
    // (i, x - c_i), which only this node's moment rows read.
    __device__ __forceinline__ void push_kept(T* __restrict__ lam_out, double* __restrict__ grad_body,
                                              const T* __restrict__ m_bar, int nn, int row, const Grid& g) const
    {
        if (grad_body) {
#pragma unroll
            for (int a = 0; a < L::D; ++a) grad_body[(size_t)a * g.n + node] = gb[a];
        }
        static_for<0, L::Q>([&](auto I) {
            constexpr int i = I;
            constexpr int c[3] = {L::cx[i], L::cy[i], L::cz[i]};
            T acc = T(0);
            acc += T(1) * m_bar[row];
#pragma unroll
            for (int a = 0; a < L::D; ++a)
                if (c[a] != 0) acc += T(c[a]) * m_bar[(size_t)(1 + a) * nn + row];
            lam_out[(size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)] = lam[i] + acc;
        });
    }
};

template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(256) void k_adj_pull(const T* __restrict__ lam_next, T* __restrict__ lam_out,
                                                  const T* __restrict__ f, const uint8_t* __restrict__ kind,
                                                  const T* __restrict__ wall_u, const T* __restrict__ field,
                                                  const T* __restrict__ body, double* __restrict__ grad_body,
                                                  T* __restrict__ nu_forced, Grid g, Relaxation<T> r)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x < g.nx && y < g.ny)
        adj_pull_node<L, T, FM, MASK>(lam_next, lam_out, f, kind, wall_u, field, body, grad_body, nu_forced, g, r, x,
                                      y, blockIdx.z);
}

// The adjoint step plus an independent link-rule pre-pass on another state buffer, in one
// launch: the grid has one extra z layer, the first, which runs the pre-pass. With
// `direct`, for a rule set in which no rule reads a slot another rule writes, each lane of
// the layer evaluates and writes one rule (as k_rules_direct), so the adjoint keeps its own
// block shape; otherwise block (0, 0) runs the two-phase pass (rules_in_one_group, at most
// 4 rules per lane). Workgroups are dispatched in block order, so the pre-pass starts with
// the first adjoint workgroups. Used to overlap the forward recompute of one checkpoint
// segment with the backward sweep of the next. Blocks may have up to 1024
// lanes (D2Q9 float32), so the pre-pass gets one lane per rule; the float64 and D3Q19
// adjoints need more than the 128 VGPRs a 1024-lane block leaves, so their blocks stay at
// 512 and 256 lanes.
constexpr int ADJ_RULES_PER_LANE = 4;
constexpr int adj_rules_max_threads(int d, size_t real) { return d == 3 ? 256 : (real == 4 ? 1024 : 512); }

template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(adj_rules_max_threads(L::D, sizeof(T))) void k_adj_pull_rules(const T* __restrict__ lam_next, T* __restrict__ lam_out,
                                                        const T* __restrict__ f, const uint8_t* __restrict__ kind,
                                                        const T* __restrict__ wall_u, const T* __restrict__ field,
                                                        const T* __restrict__ body, double* __restrict__ grad_body,
                                                        T* __restrict__ nu_forced, Grid g, Relaxation<T> r,
                                                        T* __restrict__ rf, const int64_t* __restrict__ rslot,
                                                        const T* __restrict__ rcoeff, const T* __restrict__ rconstant,
                                                        const int64_t* __restrict__ rtarget, int64_t n_rules, int rm,
                                                        int direct)
{
    if (blockIdx.z == 0) {
        if (direct) {
            const int64_t lanes = (int64_t)blockDim.x * blockDim.y;
            const int64_t rr = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * lanes + threadIdx.y * blockDim.x +
                               threadIdx.x;
            if (rr < n_rules) {
                const int64_t t = rtarget[rr];
                const T v = affine_row(rf, rslot, rcoeff, rconstant, n_rules, rm, rr);
                rf[t] = v;
            }
        } else if (blockIdx.x == 0 && blockIdx.y == 0)
            rules_in_one_group<T, ADJ_RULES_PER_LANE>(rf, rslot, rcoeff, rconstant, rtarget, n_rules, rm,
                                                      threadIdx.y * blockDim.x + threadIdx.x, blockDim.x * blockDim.y);
        return;
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x < g.nx && y < g.ny)
        adj_pull_node<L, T, FM, MASK>(lam_next, lam_out, f, kind, wall_u, field, body, grad_body, nu_forced, g, r, x,
                                      y, blockIdx.z - 1);
}

// Transposed rows: lam[slot[t]] += sum_{e in [start[t], start[t+1])} coeff[e] * x[row[e]]
// for every t < n. The slots are distinct, so no two threads touch one.
//
// A RowBatch holds, for slots t0 + k * nthreads (k < K), the slot, the list bounds and the
// first E entries of each list. fetch() reads only the transposed tables, which no launch
// writes, so it can run ahead of the barrier that publishes lam; apply() then costs one
// memory latency (the old lam values and the x reads together). Every lane reads all E
// entries of each list from an index valid either way (entry 0 past a list's end, slot 0
// past the last slot), and the values are held together, so the reads go out at once
// rather than one branch and one wait each; entries past a list's end add nothing, the
// real ones add in list order. Longer lists finish in a loop.
template <class T, int K, int E = 4>
struct RowBatch {
    int64_t sl[K];
    int b[K], end[K], nv[K];
    int rr[K][E];
    T cc[K][E];

    __device__ __forceinline__ void fetch(const int64_t* __restrict__ slot, const int* __restrict__ start,
                                          const int* __restrict__ row, const T* __restrict__ coeff, int n, int t0,
                                          int nthreads)
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int t = t0 + k * nthreads;
            const bool in = t < n;
            sl[k] = slot[in ? t : 0];
            b[k] = in ? start[t] : 0;
            end[k] = in ? start[t + 1] : 0;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            nv[k] = end[k] - b[k] < E ? end[k] - b[k] : E;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int e = j < nv[k] ? b[k] + j : 0;
                rr[k][j] = row[e];
                cc[k][j] = coeff[e];
            }
        }
    }

    __device__ __forceinline__ void apply(T* __restrict__ lam, const int* __restrict__ row,
                                          const T* __restrict__ coeff, int n, int t0, int nthreads, const T* x)
    {
        T old[K], xv[K][E];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            old[k] = lam[sl[k]];
#pragma unroll
            for (int j = 0; j < E; ++j) xv[k][j] = x[rr[k][j]];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            hold_value(old[k]);
            hold_loaded(xv[k]);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (t0 + k * nthreads < n) {
                T acc = T(0);
#pragma unroll
                for (int j = 0; j < E; ++j)
                    if (j < nv[k]) acc += cc[k][j] * xv[k][j];
                for (int e = b[k] + E; e < end[k]; ++e) acc += coeff[e] * x[row[e]];
                lam[sl[k]] = old[k] + acc;
            }
        }
    }
};

// All slots from `first` on (first = tid: every slot), K per lane at a time.
template <class T, int K = 4>
__device__ __forceinline__ void scatter_rows(T* __restrict__ lam, const int64_t* __restrict__ slot,
                                             const int* __restrict__ start, const int* __restrict__ row,
                                             const T* __restrict__ coeff, int n, const T* x, int first, int nthreads)
{
    for (int t0 = first; t0 < n; t0 += K * nthreads) {
        RowBatch<T, K> rb;
        rb.fetch(slot, start, row, coeff, n, t0, nthreads);
        rb.apply(lam, row, coeff, n, t0, nthreads, x);
    }
}

// Transposed link rules: a_r = lam[target_r]; lam[target_r] = 0; lam[slot] += sum coeff a_r
// over the (rule, coefficient) entries of each distinct term slot (host-built, fixed order).
// Rule targets are distinct, so each target is read and cleared by its own thread.
template <class T>
__global__ __launch_bounds__(1024) void k_adj_rules(T* __restrict__ lam, const int64_t* __restrict__ target,
                                                    int n_rules, const int64_t* __restrict__ uslot,
                                                    const int* __restrict__ ustart, const int* __restrict__ urule,
                                                    const T* __restrict__ ucoeff, int n_unique, T* __restrict__ a)
{
    for (int r = threadIdx.x; r < n_rules; r += blockDim.x) {
        const int64_t t = target[r];
        a[r] = lam[t];
        lam[t] = T(0);
    }
    __syncthreads();
    scatter_rows(lam, uslot, ustart, urule, ucoeff, n_unique, a, threadIdx.x, blockDim.x);
}

// ------------------------------------------------------------------ IBM adjoint

// Workgroup bound as for the forward kernel: 1024 lanes in 2D, 512 in 3D.
constexpr int ibm_adj_max_threads(int d) { return d == 2 ? 1024 : 512; }

// Forward intermediates (lbm_ibm_core.h layout, every iteration kept) followed by the
// reverse-sweep arrays.
template <class T>
struct IbmAdjScratch : IbmForward<T> {
    T *force_bar, *ustar_bar, *rho_bar, *rho_m_bar, *Fm_bar, *U_bar, *m_bar;
};

template <class T>
__host__ __device__ __forceinline__ size_t ibm_adj_layout(T* base, int d, int nn, int nm, int iters, IbmAdjScratch<T>* s)
{
    size_t o = ibm_forward_layout<T>(base, d, nn, nm, iters, s);
    auto take = [&](size_t count) {
        T* p = base + o;
        o += count;
        return p;
    };
    T* force_bar = take((size_t)d * nn);
    T* ustar_bar = take((size_t)d * nn);
    T* rho_bar = take(nn);
    T* rho_m_bar = take(nm);
    T* Fm_bar = take((size_t)d * nm);
    T* U_bar = take((size_t)d * nm);
    T* m_bar = take((size_t)(d + 1) * nn);
    if (s) {
        s->force_bar = force_bar;
        s->ustar_bar = ustar_bar;
        s->rho_bar = rho_bar;
        s->rho_m_bar = rho_m_bar;
        s->Fm_bar = Fm_bar;
        s->U_bar = U_bar;
        s->m_bar = m_bar;
    }
    return o;
}

// The reverse sweep of the forcing iterations and the moment adjoints m_bar, once the
// seeds (force_bar = nu, ustar_bar = rho_bar = rho_m_bar = 0) and the forward intermediates
// are in place behind a barrier. Touches only the workgroup's intermediates and grad_w (LDS
// when staged), never lam; ends with a barrier.
//
// Per iteration, two intervals between barriers:
//   1. marker lanes (the last lanes of the group, l = nthreads - 1 - tid): Fm_bar = total_bar
//      + sum_s w_ls force_bar(x_ls) in stencil order, then U_bar and rho_m_bar; entry lanes
//      (k = l S + s, the first lanes): grad_w[k] += sum_a Fm_a,l force_bar_a(x_k), which
//      reads force_bar before interval 2 updates it;
//   2. node lanes: u_bar from the spread lists into ustar_bar, force_bar and rho_bar; entry
//      lanes: grad_w[k] += sum_a U_bar_a,l u_a(x_k).
// The float64 weight gradients, the bulk of the arithmetic, thus run one entry per lane over
// nm 3^d lanes instead of 3^d entries in turn on nm lanes; every sum keeps its order.
template <class T, int D>
__device__ __forceinline__ void ibm_adj_reverse(const IbmAdjScratch<T> s, const IbmTables<T> tb, int nn, int nm,
                                                int iters, const T (&total_bar)[3], double* __restrict__ grad_w,
                                                int tid, int nthreads)
{
    constexpr int S = D == 2 ? 9 : 27;
    const int entries = nm * S;
    const int back = nthreads - 1 - tid;
    for (int it = iters - 1; it >= 0; --it) {
        const T* fo = s.force + (size_t)it * D * nn;
        const T* u = s.u + (size_t)it * D * nn;
        const T* U = s.U + (size_t)it * D * nm;
        const T* Fm = s.Fm + (size_t)it * D * nm;
        // Interval 1: Fm_bar, U_bar, rho_m_bar per marker.
        for (int l = back; l < nm; l += nthreads) {
            int inv[S];
            T w[S];
#pragma unroll
            for (int p = 0; p < S; ++p) {
                inv[p] = tb.inverse[(size_t)l * S + p];
                w[p] = tb.weights[(size_t)l * S + p];
            }
            T fb[3] = {total_bar[0], total_bar[1], total_bar[2]};
#pragma unroll
            for (int p = 0; p < S; ++p) {
#pragma unroll
                for (int a = 0; a < D; ++a) fb[a] += w[p] * s.force_bar[a * nn + inv[p]];
            }
            const T area = tb.areas[l];
            T rmb = T(0);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                s.Fm_bar[a * nm + l] = fb[a];
                rmb += fb[a] * T(2) * (T(0) - U[a * nm + l]) * area;
                s.U_bar[a * nm + l] = -fb[a] * T(2) * s.rho_m[l] * area;
            }
            s.rho_m_bar[l] += rmb;
        }
        // Interval 1: weight gradients of the spread, one stencil entry per lane.
        for (int k = tid; k < entries; k += nthreads) {
            const int l = k / S, x = tb.inverse[k];
            double gw = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) gw += (double)(Fm[a * nm + l] * s.force_bar[a * nn + x]);
            grad_w[k] += gw;
        }
        __syncthreads();
        // Interval 2: u_bar (one walk of the spread list for all components), then into
        // ustar_bar, force_bar and rho_bar.
        for (int n = tid; n < nn; n += nthreads) {
            T acc[D];
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] = T(0);
            for (int e = tb.spread_start[n]; e < tb.spread_start[n + 1]; ++e) {
                const int k = tb.spread_entry[e];
                const T w = tb.weights[k];
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += w * s.U_bar[a * nm + k / S];
            }
            const T rho = s.rho[n];
            const T two_rho = T(2) * rho;
            T rb = s.rho_bar[n];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                s.ustar_bar[a * nn + n] += acc[a];
                s.force_bar[a * nn + n] += acc[a] / two_rho;
                rb -= acc[a] * fo[a * nn + n] / (two_rho * rho);
            }
            s.rho_bar[n] = rb;
        }
        // Interval 2: weight gradients of the interpolation, one stencil entry per lane.
        for (int k = back; k < entries; k += nthreads) {
            const int l = k / S, x = tb.inverse[k];
            double gw = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) gw += (double)(s.U_bar[a * nm + l] * u[a * nn + x]);
            grad_w[k] += gw;
        }
        __syncthreads();
    }
    // rho_m = sum_s w rho(x_ls): weight gradient (per entry) and, per node, rho_bar and the
    // moment adjoints m_bar.
    for (int k = back; k < entries; k += nthreads) grad_w[k] += (double)(s.rho_m_bar[k / S] * s.rho[tb.inverse[k]]);
    for (int n = tid; n < nn; n += nthreads) {
        T acc = T(0);
        for (int e = tb.spread_start[n]; e < tb.spread_start[n + 1]; ++e) {
            const int k = tb.spread_entry[e];
            acc += tb.weights[k] * s.rho_m_bar[k / S];
        }
        T rb = s.rho_bar[n] + acc;
        const T rho = s.rho[n];
        for (int a = 0; a < D; ++a) {
            const T usb = s.ustar_bar[a * nn + n];
            s.m_bar[(1 + a) * nn + n] = usb / rho;
            rb -= usb * s.ustar[a * nn + n] / rho;
        }
        s.m_bar[n] = rb;
    }
    __syncthreads();
}

// One workgroup: the forward intermediates of the step (read from `record` when given,
// else recomputed from the state with the forward kernel's own code), then the iterations
// backwards. Seeds: nu (dL/dh at the forced nodes, d x n_total, written by k_adj_pull) and
// body_weight[d] = dL/d(body force). Outputs: lam[slot] += (M^T m_bar)[slot] through the
// transposed moment rows (distinct slots with their (row, coeff) lists), and
// grad_w[l][s] += dL/dw_ls in float64. With n_rules > 0 the transposed link rules are
// applied first (they act on slots the IBM rows never read).
// Byte offsets of the staged variant's LDS (16-byte aligned parts): float64 weight
// gradients, intermediates, stencil tables and, when they fit, the rule adjoints a_r.
struct AdjLdsPlan {
    size_t grad_w, base, weights, areas, inverse, entry, start, rule_a, tstart, trow, tcoeff, total;
};

// Entries of the transposed moment rows per slot: slot (k, y) is read only by the d + 1
// moment rows of the node y + c_k.
__host__ __device__ constexpr int adj_row_entries(int d) { return d + 1; }

__host__ __device__ __forceinline__ AdjLdsPlan adj_lds_plan(int d, int nn, int nm, int iters, size_t real, int lds_rules,
                                                   int lds_trows = 0)
{
    auto up = [](size_t x) { return (x + 15) & ~(size_t)15; };
    const size_t S = d == 2 ? 9 : 27;
    const size_t rules = lds_rules;
    AdjLdsPlan p{};
    size_t o = 0;
    p.grad_w = o;
    o = up(o + nm * S * sizeof(double));
    p.base = o;
    o = up(o + ibm_adj_layout<float>(nullptr, d, nn, nm, iters, nullptr) * real);
    p.weights = o;
    o = up(o + nm * S * real);
    p.areas = o;
    o = up(o + nm * real);
    p.inverse = o;
    o = up(o + nm * S * sizeof(int));
    p.entry = o;
    o = up(o + nm * S * sizeof(int));
    p.start = o;
    o = up(o + (nn + 1) * sizeof(int));
    p.rule_a = o;
    o = up(o + rules * real);
    const size_t trows = lds_trows;
    const size_t entries = trows * adj_row_entries(d);
    p.tstart = o;
    o = up(o + (trows ? trows + 1 : 0) * sizeof(int));
    p.trow = o;
    o = up(o + entries * sizeof(int));
    p.tcoeff = o;
    o = up(o + entries * real);
    p.total = o;
    return p;
}

// One workgroup: the forward intermediates of the step (read from `record` when given,
// else recomputed from the state with the forward kernel's own code), then the iterations
// backwards. Seeds: nu (dL/dh at the forced nodes, d x n_total, written by k_adj_pull) and
// body_weight[d] = dL/d(body force). Outputs: lam[slot] += (M^T m_bar)[slot] through the
// transposed moment rows (distinct slots with their (row, coeff) lists), and
// grad_w[l][s] += dL/dw_ls in float64. With n_rules > 0 the transposed link rules are
// applied first (they act on slots the IBM rows never read).
#define LBM_IBM_ADJ_PARAMS                                                                                              \
    const T *__restrict__ F, const int64_t *__restrict__ slot, const T *__restrict__ coeff,                             \
        const T *__restrict__ constant, int m, const int64_t *__restrict__ nodes, int nn,                               \
        const int *__restrict__ inverse_g, const T *__restrict__ weights_g, const T *__restrict__ areas_g, int nm,      \
        const int *__restrict__ start_g, const int *__restrict__ entry_g, int iters, T hfx, T hfy, T hfz,               \
        uint32_t n_total, const T *__restrict__ nu, const double *__restrict__ body_weight, T *__restrict__ lam,         \
        const int64_t *__restrict__ tslot, const int *__restrict__ tstart, const int *__restrict__ trow,                \
        const T *__restrict__ tcoeff, int n_tslots, double *__restrict__ grad_w_g, T *__restrict__ scratch,             \
        const int64_t *__restrict__ rtarget, int n_rules, const int64_t *__restrict__ ruslot,                           \
        const int *__restrict__ rustart, const int *__restrict__ rurule, const T *__restrict__ rucoeff, int n_runique,  \
        T *__restrict__ rule_a, int lds_rules, const T *__restrict__ record, int lds_trows
#define LBM_IBM_ADJ_ARGS                                                                                                \
    F, slot, coeff, constant, m, nodes, nn, inverse_g, weights_g, areas_g, nm, start_g, entry_g, iters, hfx, hfy, hfz,  \
        n_total, nu, body_weight, lam, tslot, tstart, trow, tcoeff, n_tslots, grad_w_g, scratch, rtarget, n_rules,      \
        ruslot, rustart, rurule, rucoeff, n_runique, rule_a, lds_rules, record, lds_trows


// Transposed slots and rule term slots whose values a lane of the prefetching variant reads
// ahead and keeps in registers across the sweep (the rest go through the generic loops).
constexpr int ADJ_AHEAD_TSLOTS = 3;
constexpr int ADJ_AHEAD_RULE_SLOTS = 2;
constexpr int ADJ_AHEAD_ENTRIES = 3;

// acc = sum over the entries [b, e) of coefficient * x[row], in list order from 0 (the
// order of RowBatch::apply).
template <class T>
__device__ __forceinline__ T row_list_sum(const int* __restrict__ row, const T* __restrict__ coeff, int b, int e,
                                          const T* __restrict__ x)
{
    T acc = T(0);
    for (int i = b; i < e; ++i) acc += coeff[i] * x[row[i]];
    return acc;
}

// The staged IBM adjoint with every global read issued before the first barrier, in two
// dependent levels: (1) target and node indices, the lane's rule term slots and transposed
// slots, the entry count, the float64 weight-gradient totals, the stencil tables and the
// record; (2) lam at the rule targets, the rule term slots and the transposed slots, nu at
// the nodes, the rule rows' entries, and the transposed rows (staged in LDS). After the
// barrier the rule adjoint, the sweep and the transposed scatter run on registers and LDS
// only, and their lam stores go out at the end. Valid when no rule term slot is a rule target
// or a transposed slot (the host's slots_disjoint flag): every value read ahead is then
// final before the kernel starts. Float32 only, with the rule adjoints a_r in LDS (float64
// values would not fit the lanes' registers). Same arithmetic in the same order as
// ibm_adj_group, so the same bits.
//
// `before` runs before the first loads and `after` once the first level has arrived, followed
// by a barrier when AFTER_BARRIER: a caller issues other loads in `before` so they share the
// first latency, and finishes work in `after` whose lam and nu stores the second level reads.
// With OWN the caller keeps the adjoint pushes of the IBM nodes (lane tid has node tid) and
// adds the transposed moment rows itself: `after(nu_seed)` gives nu at node tid, and
// `scatter(m_bar)` at the end replaces the transposed row lists, which are then neither read
// nor staged (lds_trows 0); lam at the transposed slots and nu are not read at all.
struct NoHook {
    __device__ void operator()() const {}
};

template <class T, int D, bool AFTER_BARRIER = false, bool OWN = false, class Before = NoHook, class After = NoHook,
          class Scatter = NoHook>
__device__ __forceinline__ void ibm_adj_group_prefetched(LBM_IBM_ADJ_PARAMS, int tid, int nthreads,
                                                         Before before = Before{}, After after = After{},
                                                         Scatter scatter = Scatter{})
{
    constexpr int S = D == 2 ? 9 : 27;
    constexpr int KT = ADJ_AHEAD_TSLOTS, KR = ADJ_AHEAD_RULE_SLOTS, E = ADJ_AHEAD_ENTRIES;
    extern __shared__ __align__(16) unsigned char adj_lds[];
    const AdjLdsPlan p = adj_lds_plan(D, nn, nm, iters, sizeof(T), lds_rules, lds_trows);
    double* gw = reinterpret_cast<double*>(adj_lds + p.grad_w);
    T* base = reinterpret_cast<T*>(adj_lds + p.base);
    IbmAdjScratch<T> s;
    ibm_adj_layout<T>(base, D, nn, nm, iters, &s);
    T* w = reinterpret_cast<T*>(adj_lds + p.weights);
    T* a = reinterpret_cast<T*>(adj_lds + p.areas);
    int* inv = reinterpret_cast<int*>(adj_lds + p.inverse);
    int* ent = reinterpret_cast<int*>(adj_lds + p.entry);
    int* st = reinterpret_cast<int*>(adj_lds + p.start);
    int* tst = reinterpret_cast<int*>(adj_lds + p.tstart);
    int* trw = reinterpret_cast<int*>(adj_lds + p.trow);
    T* tco = reinterpret_cast<T*>(adj_lds + p.tcoeff);
    // The rule adjoints a_r are staged too (the host picks this variant only then).
    rule_a = reinterpret_cast<T*>(adj_lds + p.rule_a);
    const IbmTables<T> tb{inv, w, a, st, ent};
    const int gw_count = nm * S;

    before();
    // ---- level 1
    const bool has_rule = tid < n_rules, has_node = tid < nn;
    const int64_t target0 = has_rule ? rtarget[tid] : 0;
    const int64_t node0 = has_node ? nodes[tid] : 0;
    int64_t usl[KR];
    int ub[KR], ue[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int t = tid + k * nthreads;
        const bool in = t < n_runique;
        usl[k] = n_runique > 0 ? ruslot[in ? t : 0] : 0;
        ub[k] = in ? rustart[t] : 0;
        ue[k] = in ? rustart[t + 1] : 0;
    }
    int64_t tsl[KT];
    int te = 0;
    if constexpr (!OWN) {
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int t = tid + k * nthreads;
            tsl[k] = n_tslots > 0 ? tslot[t < n_tslots ? t : 0] : 0;
        }
        te = tstart[n_tslots];
    }
    const double gw_old = tid < gw_count ? grad_w_g[tid] : 0.0;
    // dL/d(body force) of the step, read with the first level (the sweep needs it first).
    const T total_bar[3] = {T(-body_weight[0]), T(-body_weight[1]), T(D == 3 ? -body_weight[2] : 0.0)};
    for (int e = tid; e < gw_count; e += nthreads) gw[e] = 0.0;
    {
        const WordCopy tables[5] = {
            word_copy(inv, inverse_g, (size_t)nm * S),
            word_copy(w, weights_g, (size_t)nm * S),
            word_copy(ent, entry_g, (size_t)nm * S),
            word_copy(a, areas_g, (size_t)nm),
            word_copy(st, start_g, (size_t)nn + 1),
        };
        // Three 16-byte units per lane: the record of a 2D body goes in one round.
        const WordCopy rec[1] = {word_copy(base, record, record ? ibm_record_elems(D, nn, nm, iters) : 0)};
        copy_staged<5, 1, 3>(tables, rec, tid, nthreads);
    }
    T nu_seed[D];
    if constexpr (OWN) {
        after(nu_seed);
    } else {
        after();
    }
    if constexpr (AFTER_BARRIER) __syncthreads();

    // ---- level 2
    int rr[KR][E];
    T rc[KR][E], rold[KR];
    if (n_runique > 0) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int nv = ue[k] - ub[k] < E ? ue[k] - ub[k] : E;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int e = j < nv ? ub[k] + j : 0;
                rr[k][j] = rurule[e];
                rc[k][j] = rucoeff[e];
            }
            rold[k] = lam[usl[k]];
        }
    }
    const T lam0 = has_rule ? lam[target0] : T(0);
    T nu0[D];
    T told[KT];
    if constexpr (OWN) {
#pragma unroll
        for (int c = 0; c < D; ++c) nu0[c] = nu_seed[c];
    } else {
#pragma unroll
        for (int c = 0; c < D; ++c) nu0[c] = has_node ? nu[(size_t)c * n_total + node0] : T(0);
#pragma unroll
        for (int k = 0; k < KT; ++k) told[k] = n_tslots > 0 ? lam[tsl[k]] : T(0);
    }
    // The transposed rows go to LDS when their entries fit the plan (always, for moment rows).
    const bool rows_staged = !OWN && te <= lds_trows * adj_row_entries(D);
    if constexpr (!OWN) {
        const WordCopy none[1] = {WordCopy{nullptr, nullptr, 0}};
        const WordCopy rows[3] = {
            word_copy(tst, tstart, (size_t)n_tslots + 1),
            word_copy(trw, trow, rows_staged ? (size_t)te : 0),
            word_copy(tco, tcoeff, rows_staged ? (size_t)te : 0),
        };
        copy_staged<1, 3>(none, rows, tid, nthreads);
    }
    if (has_rule) {
        rule_a[tid] = lam0;
        lam[target0] = T(0);
    }
    if (has_node) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            s.force_bar[c * nn + tid] = nu0[c];
            s.ustar_bar[c * nn + tid] = T(0);
        }
        s.rho_bar[tid] = T(0);
    }
    for (int r = tid + nthreads; r < n_rules; r += nthreads) {
        const int64_t t = rtarget[r];
        rule_a[r] = lam[t];
        lam[t] = T(0);
    }
    for (int n = tid + nthreads; n < nn; n += nthreads) {
        for (int c = 0; c < D; ++c) {
            s.force_bar[c * nn + n] = nu[(size_t)c * n_total + nodes[n]];
            s.ustar_bar[c * nn + n] = T(0);
        }
        s.rho_bar[n] = T(0);
    }
    for (int l = tid; l < nm; l += nthreads) s.rho_m_bar[l] = T(0);
    __syncthreads();

    // ---- rule adjoint of the slots read ahead (stored at the end); later slots as usual
    T rnew[KR] = {};
    if (n_runique > 0) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < E; ++j)
                if (j < ue[k] - ub[k]) acc += rc[k][j] * rule_a[rr[k][j]];
            for (int e = ub[k] + E; e < ue[k]; ++e) acc += rucoeff[e] * rule_a[rurule[e]];
            rnew[k] = rold[k] + acc;
        }
        scatter_rows(lam, ruslot, rustart, rurule, rucoeff, n_runique, rule_a, tid + KR * nthreads, nthreads);
    }
    if (!record) {
        const int rows = (D + 1) * nn;
        const T hf[3] = {hfx, hfy, hfz};
        for (int r = tid; r < rows; r += nthreads) s.moments[r] = affine_row(F, slot, coeff, constant, rows, m, r);
        __syncthreads();
        ibm_iterate<T, D>(s, tb, nn, nm, iters, true, hf, nullptr, tid, nthreads);
    }
    ibm_adj_reverse<T, D>(s, tb, nn, nm, iters, total_bar, gw, tid, nthreads);

    // ---- transposed moment rows (from LDS when staged), then every deferred store
    const T* m_bar = s.m_bar;
    auto transposed = [&](const int* __restrict__ rw, const T* __restrict__ cf) {
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int t = tid + k * nthreads;
            if (t < n_tslots) lam[tsl[k]] = told[k] + row_list_sum(rw, cf, tst[t], tst[t + 1], m_bar);
        }
        for (int t = tid + KT * nthreads; t < n_tslots; t += nthreads)
            lam[tslot[t]] = lam[tslot[t]] + row_list_sum(rw, cf, tst[t], tst[t + 1], m_bar);
    };
    if constexpr (OWN) {
        scatter(m_bar);
    } else {
        if (rows_staged) transposed(trw, tco);
        else transposed(trow, tcoeff);
    }
    if (n_runique > 0) {
#pragma unroll
        for (int k = 0; k < KR; ++k)
            if (tid + k * nthreads < n_runique) lam[usl[k]] = rnew[k];
    }
    if (tid < gw_count) grad_w_g[tid] = gw_old + gw[tid];
    for (int e = tid + nthreads; e < gw_count; e += nthreads) grad_w_g[e] += gw[e];
}

template <class T, int D, bool STAGED>
__device__ __forceinline__ void ibm_adj_group(LBM_IBM_ADJ_PARAMS, int tid, int nthreads)
{
    if constexpr (STAGED && sizeof(T) == 4) {
        if (lds_trows > 0) {
            ibm_adj_group_prefetched<T, D>(LBM_IBM_ADJ_ARGS, tid, nthreads);
            return;
        }
    }
    constexpr int S = D == 2 ? 9 : 27;
    extern __shared__ __align__(16) unsigned char adj_lds[];
    // Index loads of the first interval go out before the table copies, so their second
    // level (lam at the rule targets, nu at the nodes) costs one more latency, not three.
    const bool has_rule = tid < n_rules, has_node = tid < nn;
    const int64_t target0 = has_rule ? rtarget[tid] : 0;
    const int64_t node0 = has_node ? nodes[tid] : 0;
    // The first two transposed rule slots of this lane, fetched ahead of the barrier.
    RowBatch<T, 2> rule_rows;
    if (n_rules > 0) rule_rows.fetch(ruslot, rustart, rurule, rucoeff, n_runique, tid, nthreads);
    IbmAdjScratch<T> s;
    IbmTables<T> tb{inverse_g, weights_g, areas_g, start_g, entry_g};
    double* grad_w = grad_w_g;
    T* base;
    if constexpr (STAGED) {
        const AdjLdsPlan p = adj_lds_plan(D, nn, nm, iters, sizeof(T), lds_rules);
        double* gw = reinterpret_cast<double*>(adj_lds + p.grad_w);
        base = reinterpret_cast<T*>(adj_lds + p.base);
        ibm_adj_layout<T>(base, D, nn, nm, iters, &s);
        T* w = reinterpret_cast<T*>(adj_lds + p.weights);
        T* a = reinterpret_cast<T*>(adj_lds + p.areas);
        int* inv = reinterpret_cast<int*>(adj_lds + p.inverse);
        int* ent = reinterpret_cast<int*>(adj_lds + p.entry);
        int* st = reinterpret_cast<int*>(adj_lds + p.start);
        if (lds_rules) rule_a = reinterpret_cast<T*>(adj_lds + p.rule_a);
        for (int e = tid; e < nm * S; e += nthreads) gw[e] = 0.0;
        const WordCopy tables[5] = {
            word_copy(inv, inverse_g, (size_t)nm * S),
            word_copy(w, weights_g, (size_t)nm * S),
            word_copy(ent, entry_g, (size_t)nm * S),
            word_copy(a, areas_g, (size_t)nm),
            word_copy(st, start_g, (size_t)nn + 1),
        };
        copy_tables(tables, word_copy(base, record, record ? ibm_record_elems(D, nn, nm, iters) : 0), tid, nthreads);
        tb = IbmTables<T>{inv, w, a, st, ent};
        grad_w = gw;
    } else {
        base = scratch;
        ibm_adj_layout<T>(scratch, D, nn, nm, iters, &s);
        if (record) {
            const WordCopy none[1] = {WordCopy{nullptr, nullptr, 0}};
            copy_tables(none, word_copy(base, record, ibm_record_elems(D, nn, nm, iters)), tid, nthreads);
        }
    }
    // First interval, all independent: the tables above, the recorded intermediates, the
    // rule adjoints a_r (targets cleared) and the reverse-sweep seeds.
    {
        const T lam0 = has_rule ? lam[target0] : T(0);
        T nu0[D];
#pragma unroll
        for (int a = 0; a < D; ++a) nu0[a] = has_node ? nu[(size_t)a * n_total + node0] : T(0);
        if (has_rule) {
            rule_a[tid] = lam0;
            lam[target0] = T(0);
        }
        if (has_node) {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                s.force_bar[a * nn + tid] = nu0[a];
                s.ustar_bar[a * nn + tid] = T(0);
            }
            s.rho_bar[tid] = T(0);
        }
    }
    for (int r = tid + nthreads; r < n_rules; r += nthreads) {
        const int64_t t = rtarget[r];
        rule_a[r] = lam[t];
        lam[t] = T(0);
    }
    for (int n = tid + nthreads; n < nn; n += nthreads) {
        for (int a = 0; a < D; ++a) {
            s.force_bar[a * nn + n] = nu[(size_t)a * n_total + nodes[n]];
            s.ustar_bar[a * nn + n] = T(0);
        }
        s.rho_bar[n] = T(0);
    }
    for (int l = tid; l < nm; l += nthreads) s.rho_m_bar[l] = T(0);
    __syncthreads();
    if (n_rules > 0) {
        rule_rows.apply(lam, rurule, rucoeff, n_runique, tid, nthreads, rule_a);
        scatter_rows(lam, ruslot, rustart, rurule, rucoeff, n_runique, rule_a, tid + 2 * nthreads, nthreads);
    }
    if (!record) {
        const int rows = (D + 1) * nn;
        const T hf[3] = {hfx, hfy, hfz};
        for (int r = tid; r < rows; r += nthreads) s.moments[r] = affine_row(F, slot, coeff, constant, rows, m, r);
        __syncthreads();
        ibm_iterate<T, D>(s, tb, nn, nm, iters, true, hf, nullptr, tid, nthreads);
    }
    const T total_bar[3] = {T(-body_weight[0]), T(-body_weight[1]), T(D == 3 ? -body_weight[2] : 0.0)};
    ibm_adj_reverse<T, D>(s, tb, nn, nm, iters, total_bar, grad_w, tid, nthreads);
    scatter_rows(lam, tslot, tstart, trow, tcoeff, n_tslots, s.m_bar, tid, nthreads);
    if constexpr (STAGED) {
        for (int e = tid; e < nm * S; e += nthreads) grad_w_g[e] += grad_w[e];
    }
}

template <class T, int D, bool STAGED>
__global__ __launch_bounds__(ibm_adj_max_threads(D)) void k_ibm_adj(LBM_IBM_ADJ_PARAMS)
{
    ibm_adj_group<T, D, STAGED>(LBM_IBM_ADJ_ARGS, threadIdx.x, blockDim.x);
}

// The IBM adjoint (block 0) and an independent pull step on other buffers (blocks 1..,
// tiles of pbx x pby nodes, lanes beyond pbx * pby idle) in one launch. Used to overlap the
// forward recompute of one checkpoint segment with the backward sweep of the next; the
// long single-group IBM adjoint is dispatched first.
template <class L, class T, ForceMode FM, bool MASK, bool STAGED>
__global__ __launch_bounds__(ibm_adj_max_threads(L::D)) void k_ibm_adj_pull(
    LBM_IBM_ADJ_PARAMS, const T* __restrict__ src, T* __restrict__ dst, const uint8_t* __restrict__ kind,
    const T* __restrict__ wall_u, const T* __restrict__ field, const T* __restrict__ body, Grid g, Relaxation<T> r,
    int pbx, int pby, int tiles_x, int tiles_y)
{
    if (blockIdx.x == 0) {
        ibm_adj_group<T, L::D, STAGED>(LBM_IBM_ADJ_ARGS, threadIdx.x, blockDim.x);
        return;
    }
    const int t = blockIdx.x - 1;
    const int lane = threadIdx.x;
    if (lane >= pbx * pby) return;
    const int tx = t % tiles_x, rest = t / tiles_x;
    Coord c{(tx * pbx) + lane % pbx, (rest % tiles_y) * pby + lane / pbx, rest / tiles_y};
    if (c.x < g.nx && c.y < g.ny) pull_update<L, T, FM, MASK>(src, dst, kind, wall_u, field, body, g, r, c);
}

// ------------------------------------------------------------------ fused backward step

// What the IBM workgroup of the next backward step reads first: the state it pulls from,
// its record and its field (null when there is no next step in the same sweep), and the
// tile workgroup (`block`, -1 for none) that reads them ahead.
template <class T>
struct AdjNext {
    const T* f;
    const T* record;
    const T* field;
    int block;
};

// Reads, once its own nodes are done, what the next step's IBM workgroup will read first
// (the region list; kind, body force, field and pulled populations of the region nodes; the
// record; the stencil tables), so that workgroup finds them in a cache (the L2 of its XCD,
// the host picking a block on the same XCD, or the Infinity Cache all XCDs share) instead of
// HBM. None of it is written in this launch. The values are summed and held, so the loads
// are issued; nothing is written.
template <class L, class T>
__device__ __forceinline__ void adj_prefetch_next(const AdjNext<T>& next, const int32_t* __restrict__ region,
                                                  int n_region, const uint8_t* __restrict__ kind,
                                                  const T* __restrict__ body, const int* __restrict__ inverse,
                                                  const T* __restrict__ weights, const T* __restrict__ areas,
                                                  const int* __restrict__ start, const int* __restrict__ entry,
                                                  int nn, int nm, int iters, const Grid& g, int tid, int lanes)
{
    constexpr int S = L::D == 2 ? 9 : 27;
    uint32_t sum = 0;
    for (int q = tid; q < n_region; q += lanes) {
        const uint32_t node = (uint32_t)region[q];
        const uint32_t rest = node / (uint32_t)g.nx;
        const int x = (int)(node - rest * (uint32_t)g.nx), y = (int)(rest % (uint32_t)g.ny),
                  z = (int)(rest / (uint32_t)g.ny);
        T v = T(0);
        static_for<0, L::Q>([&](auto I) { v = v + next.f[(size_t)I * g.n + nbr<L, I, -1>(x, y, z, g)]; });
#pragma unroll
        for (int a = 0; a < L::D; ++a) v = v + body[(size_t)a * g.n + node] + next.field[(size_t)a * g.n + node];
        sum += (uint32_t)kind[node] + (v != v ? 1u : 0u);
    }
    const size_t rec = ibm_record_elems(L::D, nn, nm, iters);
    T v = T(0);
    // Eight loads in flight per lane and pass (clamped indices), then their sum.
    for (size_t w0 = tid; w0 < rec; w0 += (size_t)8 * lanes) {
        T a[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const size_t w = w0 + (size_t)k * lanes;
            a[k] = next.record[w < rec ? w : rec - 1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) v = v + a[k];
    }
    for (int w = tid; w < nm * S; w += lanes) v = v + weights[w] + T(inverse[w] + entry[w]);
    for (int w = tid; w < nm; w += lanes) v = v + areas[w];
    for (int w = tid; w <= nn; w += lanes) sum += (uint32_t)start[w];
    sum += v != v ? 1u : 0u;
    hold_value(sum);
}

// The link-rule nodes of the fused backward step, for rule sets in which every slot a rule
// reads or writes is written by the adjoint step of one node, the rule's node: info[k] =
// {node, directions i whose upstream node x - c_i is solid | the node's rule target
// directions << 16, 0, 0}, where rule target direction j is the push of the node that writes
// the rule's target slot; coeff[k][i][j] (Q x Q) = the coefficient of the rule with target
// direction j in the transposed row of the slot push i writes (zero where the rule does not
// read that slot). Nodes sorted by index, in chunks of `chunk` nodes.
template <class T>
struct AdjRuleNodes {
    const uint4* info;
    const T* coeff;
    int count;
    int chunk;
};

// One chunk of the link-rule nodes by one workgroup: the adjoint step of each node with the
// transposed link rules applied to its pushes before they are stored. As lbm_adj_rules after
// the pushes: a_j = the push of target direction j, that push becomes zero, and push i gets
// sum_j coeff[i][j] a_j over the nonzero coefficients in increasing j (the transposed row's
// rule order: a node's rules are sorted by direction), added from 0 and only where the row
// has entries, so the same bits.
template <class L, class T, ForceMode FM>
__device__ __forceinline__ void adj_rule_chunk(const T* __restrict__ lam_next, T* __restrict__ lam,
                                               const T* __restrict__ f, const uint8_t* __restrict__ kind,
                                               const T* __restrict__ wall_u, const T* __restrict__ field,
                                               const T* __restrict__ body, double* __restrict__ grad_body,
                                               T* __restrict__ nu_forced, const Grid& g, const Relaxation<T>& r,
                                               const AdjRuleNodes<T>& ar, int c, int lanes, int tid,
                                               T* __restrict__ cl)
{
    constexpr int Q = L::Q;
    const int first = c * ar.chunk;
    const int end = min(first + ar.chunk, ar.count);
    const bool node_lane = first + tid < end;
    uint4 info = make_uint4(0u, 0u, 0u, 0u);
    if (node_lane) info = ar.info[first + tid];
    AdjNodeParts<L, T, FM, true> nd;
    if (node_lane) nd.load(lam_next, f, kind, wall_u, field, body, grad_body, g, info.x, info.y & 0xffffu);
    {
        const WordCopy none[1] = {WordCopy{nullptr, nullptr, 0}};
        const WordCopy cm[1] = {word_copy(cl, ar.coeff + (size_t)first * Q * Q, (size_t)(end - first) * Q * Q)};
        copy_staged<1, 1>(cm, none, tid, lanes);
    }
    __syncthreads();
    if (!node_lane) return;
    T mu[Q];
    T Fbx, Fby, Fbz;
    nd.adjoint(wall_u, r, mu, Fbx, Fby, Fbz);
    const uint32_t targets = info.y >> 16;
    const T* cm = cl + (size_t)tid * Q * Q;
    // a_j for every direction (only the targets' are read); a select between the pushes and
    // a zero would make the compiler select addresses and keep mu in scratch memory.
    T a[Q];
    static_for<0, Q>([&](auto J) {
        a[J] = mu[J];
        if ((targets >> J) & 1u) mu[J] = T(0);
    });
    static_for<0, Q>([&](auto I) {
        T acc = T(0);
        bool any = false;
        static_for<0, Q>([&](auto J) {
            if ((targets >> J) & 1u) {
                const T cij = cm[(int)I * Q + (int)J];
                if (cij != T(0)) {
                    acc += cij * a[J];
                    any = true;
                }
            }
        });
        if (any) mu[I] = mu[I] + acc;
    });
    nd.store(lam, mu, grad_body, nu_forced, g, Fbx, Fby, Fbz);
}

//
// One backward step of the pipelined sweep in one launch (float32, the prefetching IBM
// adjoint; a one-dimensional grid of workgroups of bx x by <= 1024 lanes):
//   workgroup 0        the IBM region: the adjoint step of the IBM stencil nodes (listed in
//                      `region`; none has a solid neighbour), their loads issued with the
//                      IBM adjoint's first level, a barrier, then the rest of the IBM adjoint
//                      (ibm_adj_group_prefetched without rules), whose lam reads find values
//                      this workgroup has just written; with `keep` the stencil nodes keep
//                      their pushes and nu in registers instead and add their moment rows'
//                      adjoint at the end;
//   next `achunks`     the link-rule nodes (adj_rule_chunk): their adjoint step with the
//                      transposed link rules applied to their pushes;
//   next `chunks`      the partner (forward recompute) step's rule chunks, with write-back;
//   next ptiles        the partner step's other nodes (pull_update, rule nodes left out);
//   last atiles        the adjoint step of every node outside the regions.
// Every lam slot of a fluid node is written by exactly one node's push; the slots the IBM
// adjoint reads or updates are written by IBM-region nodes and those of the link rules by the
// rules' own nodes (the host keeps the IBM region, the rule nodes and their slot sets apart),
// so no workgroup reads a slot another one writes; the partner step works on other buffers.
// The operations are those of the two-launch form (adjoint step with its riding pre-pass,
// then the IBM and rule adjoints with the riding pull), so the bits are the same.
template <class L, class T, ForceMode FM>
__global__ __launch_bounds__(1024) void k_adj_fused(LBM_IBM_ADJ_PARAMS, const T* __restrict__ lam_next,
                                                    const uint8_t* __restrict__ kind, const T* __restrict__ wall_u,
                                                    const T* __restrict__ field, const T* __restrict__ body,
                                                    double* __restrict__ grad_body,
                                                    const int32_t* __restrict__ region, int n_region, int keep,
                                                    AdjRuleNodes<T> ar,
                                                    Grid g, Relaxation<T> r, int atiles_x, int atiles_y,
                                                    const T* __restrict__ p_src, T* __restrict__ p_dst,
                                                    const T* __restrict__ p_field, RuleRows<T> p_rules, int ptiles_x,
                                                    int ptiles_y, int ptiles, AdjNext<T> next)
{
    const int b = (int)blockIdx.x;
    const int lanes = (int)(blockDim.x * blockDim.y);
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    T* nu_out = const_cast<T*>(nu);  // the IBM region writes nu, then the IBM adjoint reads it
    if (b == 0) {
        // The stencil nodes' adjoint step: the first `lanes` nodes load alongside the IBM
        // adjoint's first level and finish once it has arrived; a barrier, then the IBM
        // adjoint alone reads their lam and nu (workgroup 1 runs the transposed link rules).
        AdjNodeParts<L, T, FM> own;
        const bool has = tid < n_region;
        auto before = [&] {
            if (has) own.load(lam_next, F, kind, wall_u, field, body, grad_body, g, (uint32_t)region[tid]);
        };
        auto after = [&] {
            if (has) own.finish(lam, grad_body, nu_out, wall_u, g, r);
            for (int q = tid + lanes; q < n_region; q += lanes) {
                const uint32_t node = (uint32_t)region[q];
                const uint32_t rest = node / (uint32_t)g.nx;
                adj_pull_node<L, T, FM, true>(lam_next, lam, F, kind, wall_u, field, body, grad_body, nu_out, g, r,
                                              (int)(node - rest * (uint32_t)g.nx), (int)(rest % (uint32_t)g.ny),
                                              (int)(rest / (uint32_t)g.ny));
            }
        };
        if (keep) {
            // The region is the IBM node list itself (node tid is IBM node tid) and its pushes
            // are the transposed slots: each node keeps its pushes and nu in registers and adds
            // its moment rows' adjoint at the end. No barrier after the first level: the rest
            // reads nothing these nodes would have written.
            auto after_kept = [&](T (&nu_seed)[L::D]) {
#pragma unroll
                for (int a = 0; a < L::D; ++a) nu_seed[a] = T(0);
                if (has) own.finish_kept(nu_seed, r);
            };
            auto scatter_kept = [&](const T* m_bar) {
                if (has) own.push_kept(lam, grad_body, m_bar, nn, tid, g);
            };
            ibm_adj_group_prefetched<T, L::D, false, true>(
                F, slot, coeff, constant, m, nodes, nn, inverse_g, weights_g, areas_g, nm, start_g, entry_g, iters, hfx,
                hfy, hfz, n_total, nu, body_weight, lam, tslot, tstart, trow, tcoeff, n_tslots, grad_w_g, scratch,
                nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr, 0, record, 0, tid, lanes, before,
                after_kept, scatter_kept);
            return;
        }
        ibm_adj_group_prefetched<T, L::D, true>(F, slot, coeff, constant, m, nodes, nn, inverse_g, weights_g, areas_g,
                                                nm, start_g, entry_g, iters, hfx, hfy, hfz, n_total, nu, body_weight,
                                                lam, tslot, tstart, trow, tcoeff, n_tslots, grad_w_g, scratch,
                                                nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr, 0, record,
                                                lds_trows, tid, lanes, before, after);
        return;
    }
    int t = b - 1;
    const int achunks = ar.count ? rule_chunk_count(ar.count, ar.chunk) : 0;
    if (t < achunks) {
        extern __shared__ __align__(16) unsigned char adj_lds[];
        adj_rule_chunk<L, T, FM>(lam_next, lam, F, kind, wall_u, field, body, grad_body, nu_out, g, r, ar, t, lanes,
                                 tid, reinterpret_cast<T*>(adj_lds));
        return;
    }
    t -= achunks;
    const int chunks = p_rules.count ? rule_chunk_count(p_rules.count, p_rules.chunk) : 0;
    if (t < chunks) {
        extern __shared__ double dynamic_lds[];
        rule_chunk<L, T, FM, true>(p_src, p_dst, kind, wall_u, p_field, body, g, r, p_rules, t, lanes, tid,
                                   reinterpret_cast<T*>(dynamic_lds));
        return;
    }
    t -= chunks;
    if (t < ptiles) {
        Coord c;
        c.x = (t % ptiles_x) * (int)blockDim.x + (int)threadIdx.x;
        c.y = ((t / ptiles_x) % ptiles_y) * (int)blockDim.y + (int)threadIdx.y;
        c.z = t / (ptiles_x * ptiles_y);
        if (c.x < g.nx && c.y < g.ny) pull_update<L, T, FM, true, true>(p_src, p_dst, kind, wall_u, p_field, body, g, r, c);
        return;
    }
    t -= ptiles;
    const int x = (t % atiles_x) * (int)blockDim.x + (int)threadIdx.x;
    const int y = ((t / atiles_x) % atiles_y) * (int)blockDim.y + (int)threadIdx.y;
    const int z = t / (atiles_x * atiles_y);
    if (x < g.nx && y < g.ny)
        adj_pull_node<L, T, FM, true, true>(lam_next, lam, F, kind, wall_u, field, body, grad_body, nu_out, g, r, x,
                                            y, z);
    if (b == next.block)
        adj_prefetch_next<L, T>(next, region, n_region, kind, body, inverse_g, weights_g, areas_g, start_g, entry_g,
                                nn, nm, iters, g, tid, lanes);
}

#undef LBM_IBM_ADJ_PARAMS
#undef LBM_IBM_ADJ_ARGS

// Call go(FM, MASK) for the force and mask modes of the C ABI (field forcing needs the kind
// mask).
template <class Go>
void dispatch_force_mask(int force_mode, int mask_mode, Go&& go)
{
    using U = std::integral_constant<ForceMode, ForceMode::UNIFORM>;
    using Fi = std::integral_constant<ForceMode, ForceMode::FIELD>;
    using De = std::integral_constant<ForceMode, ForceMode::DENSE>;
    using M0 = std::integral_constant<bool, false>;
    using M1 = std::integral_constant<bool, true>;
    if (force_mode == 0) {
        if (mask_mode) go(U{}, M1{});
        else go(U{}, M0{});
    } else if (force_mode == 1) {
        go(Fi{}, M1{});
    } else {
        if (mask_mode) go(De{}, M1{});
        else go(De{}, M0{});
    }
}

void check_step_args(int force_mode, int mask_mode, const void* kind, const void* wall_u, const void* force_field,
                     const void* body_field)
{
    require(mask_mode == 0 || kind, "mask_mode 1 needs the kind array");
    require(force_mode >= 0 && force_mode <= 2, "unknown force mode");
    require(force_mode != 1 || (force_field && mask_mode == 1), "field force mode needs the field and kind");
    require(force_mode != 2 || body_field, "dense force mode needs a body field");
}

// Launch shape of the IBM adjoint: lanes, LDS bytes, staging, rule-adjoint placement and
// (the prefetching variant) the transposed rows staged in LDS.
struct IbmAdjShape {
    int threads;
    size_t lds;
    bool staged;
    int rules_in_lds;
    int rows_in_lds;
};

template <class L, class T>
IbmAdjShape ibm_adj_shape(int n_nodes, int n_markers, int iterations, int n_rules, int n_tslots, bool disjoint)
{
    const size_t limit = lds_limit();
    const size_t lds_tables = adj_lds_plan(L::D, n_nodes, n_markers, iterations, sizeof(T), 0).total;
    const size_t lds_rules = adj_lds_plan(L::D, n_nodes, n_markers, iterations, sizeof(T), n_rules).total;
    IbmAdjShape sh;
    sh.staged = lds_tables <= limit;
    sh.rules_in_lds = sh.staged && n_rules > 0 && lds_rules <= limit ? n_rules : 0;
    sh.lds = sh.staged ? (sh.rules_in_lds ? lds_rules : lds_tables) : 0;
    // The prefetching variant (float32): staged with the rule adjoints in LDS, slots
    // disjoint, and room for the transposed rows.
    const size_t lds_rows =
        adj_lds_plan(L::D, n_nodes, n_markers, iterations, sizeof(T), sh.rules_in_lds, n_tslots).total;
    sh.rows_in_lds = sizeof(T) == 4 && sh.staged && disjoint && n_tslots > 0 && (n_rules == 0 || sh.rules_in_lds) &&
                             lds_rows <= limit
                         ? n_tslots
                         : 0;
    if (sh.rows_in_lds) sh.lds = lds_rows;
    const int rows = (L::D + 1) * n_nodes;
    sh.threads = std::max(256, std::min(ibm_adj_max_threads(L::D), 64 * ((rows + 63) / 64)));
    return sh;
}

void check_ibm_adj_args(int m, int n_nodes, int n_markers, int iterations, int64_t n_total, int n_rules,
                        std::initializer_list<const void*> required, std::initializer_list<const void*> rule_ptrs)
{
    require(n_nodes > 0 && n_markers > 0 && iterations >= 1 && m >= 1, "bad IBM sizes");
    require(n_total > 0 && n_total < (1ll << 31), "node count out of range");
    for (const void* p : required) require(p != nullptr, "null pointer");
    if (n_rules > 0)
        for (const void* p : rule_ptrs) require(p != nullptr, "null rule pointer");
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LBM_API int lbm_adj_pull_step(int lattice, int dtype, const void* lam_next, void* lam_out, const void* f,
                              const uint8_t* kind, const void* wall_u, const void* force_field, const void* body_field,
                              int nx, int ny, int nz, double omega_plus, double omega_minus, double fx, double fy,
                              double fz, int force_mode, int mask_mode, double* grad_body, void* nu_forced, int bx,
                              int by, void* stream)
{
    return guarded("lbm_adj_pull_step", [&] {
        const Grid g = make_grid(lattice, nx, ny, nz);
        require(bx > 0 && by > 0 && bx * by <= 256, "adjoint block must be positive with at most 256 threads");
        require(lam_next && lam_out && f && lam_next != lam_out, "adjoint buffers must be distinct and non-null");
        check_step_args(force_mode, mask_mode, kind, wall_u, force_field, body_field);
        const double force[3] = {fx, fy, fz};
        const dim3 block(bx, by, 1);
        const dim3 grid = grid_of(g, block);
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const Relaxation<T> r = make_relaxation<L, T>(omega_plus, omega_minus, force);
            dispatch_force_mask(force_mode, mask_mode, [&](auto fm, auto mk) {
                k_adj_pull<L, T, decltype(fm)::value, decltype(mk)::value><<<grid, block, 0, as_stream(stream)>>>(
                    static_cast<const T*>(lam_next), static_cast<T*>(lam_out), static_cast<const T*>(f), kind,
                    static_cast<const T*>(wall_u), static_cast<const T*>(force_field),
                    static_cast<const T*>(body_field), grad_body, static_cast<T*>(nu_forced), g, r);
            });
            check_launch("adjoint step");
        });
    });
}

LBM_API int lbm_adj_pull_step_rules(int lattice, int dtype, const void* lam_next, void* lam_out, const void* f,
                                    const uint8_t* kind, const void* wall_u, const void* force_field,
                                    const void* body_field, int nx, int ny, int nz, double omega_plus,
                                    double omega_minus, double fx, double fy, double fz, int force_mode,
                                    int mask_mode, double* grad_body, void* nu_forced, int bx, int by, void* rule_f,
                                    const int64_t* rule_slot, const void* rule_coeff, const void* rule_constant,
                                    const int64_t* rule_target, int64_t n_rules, int rule_m, int rules_direct,
                                    void* stream)
{
    return guarded("lbm_adj_pull_step_rules", [&] {
        const Grid g = make_grid(lattice, nx, ny, nz);
        const int max_threads = adj_rules_max_threads(lattice == LBM_D2Q9 ? 2 : 3, dtype == LBM_F32 ? 4 : 8);
        require(bx > 0 && by > 0 && bx * by <= max_threads,
                "block must be positive with at most 1024 (D2Q9 f32), 512 (D2Q9 f64) or 256 (D3Q19) threads");
        require(lam_next && lam_out && f && lam_next != lam_out, "adjoint buffers must be distinct and non-null");
        check_step_args(force_mode, mask_mode, kind, wall_u, force_field, body_field);
        const dim3 layer = grid_of(g, dim3(bx, by, 1));
        const int64_t layer_lanes = (int64_t)layer.x * layer.y * bx * by;
        require(n_rules >= 1 && n_rules <= (rules_direct ? layer_lanes : (int64_t)ADJ_RULES_PER_LANE * bx * by),
                "the fused pre-pass takes 1 .. 4 * bx * by rules, or one per lane of a grid layer when direct");
        require(rule_m >= 1 && rule_m <= 128, "rules need 1 <= m <= 128 terms");
        require(rule_f && rule_slot && rule_coeff && rule_constant && rule_target, "null rule pointer");
        require(rule_f != f && rule_f != lam_out, "the pre-pass state must be another buffer");
        const double force[3] = {fx, fy, fz};
        const dim3 block(bx, by, 1);
        dim3 grid = grid_of(g, block);
        grid.z += 1;  // the pre-pass layer (z = 0)
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const Relaxation<T> r = make_relaxation<L, T>(omega_plus, omega_minus, force);
            dispatch_force_mask(force_mode, mask_mode, [&](auto fm, auto mk) {
                k_adj_pull_rules<L, T, decltype(fm)::value, decltype(mk)::value>
                    <<<grid, block, 0, as_stream(stream)>>>(
                        static_cast<const T*>(lam_next), static_cast<T*>(lam_out), static_cast<const T*>(f), kind,
                        static_cast<const T*>(wall_u), static_cast<const T*>(force_field),
                        static_cast<const T*>(body_field), grad_body, static_cast<T*>(nu_forced), g, r,
                        static_cast<T*>(rule_f), rule_slot, static_cast<const T*>(rule_coeff),
                        static_cast<const T*>(rule_constant), rule_target, n_rules, rule_m, rules_direct);
            });
            check_launch("adjoint step with pre-pass");
        });
    });
}

LBM_API int lbm_adj_rules(int dtype, void* lam, const int64_t* target, int n_rules, const int64_t* uslot,
                          const int* ustart, const int* urule, const void* ucoeff, int n_unique, void* scratch,
                          void* stream)
{
    return guarded("lbm_adj_rules", [&] {
        if (n_rules == 0) return;
        require(lam && target && uslot && ustart && urule && ucoeff && scratch, "null pointer");
        dispatch_real(dtype, [&](auto Tt) {
            using T = decltype(Tt);
            k_adj_rules<T><<<1, 1024, 0, as_stream(stream)>>>(static_cast<T*>(lam), target, n_rules, uslot, ustart,
                                                              urule, static_cast<const T*>(ucoeff), n_unique,
                                                              static_cast<T*>(scratch));
            check_launch("adjoint rules");
        });
    });
}

LBM_API int lbm_ibm_adjoint_scratch_size(int lattice, int dtype, int n_nodes, int n_markers, int iterations,
                                         int64_t* bytes)
{
    return guarded("lbm_ibm_adjoint_scratch_size", [&] {
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            *bytes = (int64_t)(ibm_adj_layout<T>(nullptr, L::D, n_nodes, n_markers, iterations, nullptr) * sizeof(T));
        });
    });
}

#define LBM_IBM_ADJ_CALL_ARGS                                                                                           \
    static_cast<const T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), m, nodes, n_nodes,    \
        inverse, static_cast<const T*>(weights), static_cast<const T*>(areas), n_markers, spread_start, spread_entry,   \
        iterations, T(0.5 * fx), T(0.5 * fy), T(0.5 * fz), (uint32_t)n_total, static_cast<const T*>(nu), body_weight,  \
        static_cast<T*>(lam), tslot, tstart, trow, static_cast<const T*>(tcoeff), n_tslots, grad_w,                     \
        static_cast<T*>(scratch), rule_target, n_rules, rule_uslot, rule_ustart, rule_urule,                           \
        static_cast<const T*>(rule_ucoeff), n_rule_unique, static_cast<T*>(rule_scratch), sh.rules_in_lds,             \
        static_cast<const T*>(record), sh.rows_in_lds

#define LBM_IBM_ADJ_CHECK                                                                                              \
    check_ibm_adj_args(m, n_nodes, n_markers, iterations, n_total, n_rules,                                            \
                       {f, slot, coeff, constant, nodes, inverse, weights, areas, spread_start, spread_entry, nu,      \
                        body_weight, lam, tslot, tstart, trow, tcoeff, grad_w, scratch},                               \
                       {rule_target, rule_uslot, rule_ustart, rule_urule, rule_ucoeff, rule_scratch})

LBM_API int lbm_ibm_adjoint(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                            const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                            const void* weights, const void* areas, int n_markers, const int* spread_start,
                            const int* spread_entry, int iterations, double fx, double fy, double fz, int64_t n_total,
                            const void* nu, const double* body_weight, void* lam, const int64_t* tslot,
                            const int* tstart, const int* trow, const void* tcoeff, int n_tslots, double* grad_w,
                            void* scratch, const int64_t* rule_target, int n_rules, const int64_t* rule_uslot,
                            const int* rule_ustart, const int* rule_urule, const void* rule_ucoeff, int n_rule_unique,
                            void* rule_scratch, int slots_disjoint, const void* record, void* stream)
{
    return guarded("lbm_ibm_adjoint", [&] {
        LBM_IBM_ADJ_CHECK;
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const IbmAdjShape sh =
                ibm_adj_shape<L, T>(n_nodes, n_markers, iterations, n_rules, n_tslots, slots_disjoint != 0);
            auto go = [&](auto staged) {
                constexpr bool ST = decltype(staged)::value;
                if (ST) allow_lds(k_ibm_adj<T, L::D, ST>, sh.lds);
                k_ibm_adj<T, L::D, ST><<<1, sh.threads, sh.lds, as_stream(stream)>>>(LBM_IBM_ADJ_CALL_ARGS);
            };
            if (sh.staged) go(std::true_type{});
            else go(std::false_type{});
            check_launch("IBM adjoint");
        });
    });
}

LBM_API int lbm_ibm_adjoint_pull(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                                 const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                                 const void* weights, const void* areas, int n_markers, const int* spread_start,
                                 const int* spread_entry, int iterations, double fx, double fy, double fz,
                                 int64_t n_total, const void* nu, const double* body_weight, void* lam,
                                 const int64_t* tslot, const int* tstart, const int* trow, const void* tcoeff,
                                 int n_tslots, double* grad_w, void* scratch, const int64_t* rule_target, int n_rules,
                                 const int64_t* rule_uslot, const int* rule_ustart, const int* rule_urule,
                                 const void* rule_ucoeff, int n_rule_unique, void* rule_scratch, int slots_disjoint,
                                 const void* record, const void* pull_src, void* pull_dst, const uint8_t* kind, const void* wall_u,
                                 const void* force_field, const void* body_field, int nx, int ny, int nz,
                                 double omega_plus, double omega_minus, int force_mode, int mask_mode, void* stream)
{
    return guarded("lbm_ibm_adjoint_pull", [&] {
        LBM_IBM_ADJ_CHECK;
        const Grid g = make_grid(lattice, nx, ny, nz);
        require((int64_t)g.n == n_total, "the pull grid must have n_total nodes");
        require(pull_src && pull_dst && pull_src != pull_dst, "pull buffers must be distinct and non-null");
        require(pull_src != f && pull_dst != f && pull_dst != lam, "the pull must run on other buffers");
        check_step_args(force_mode, mask_mode, kind, wall_u, force_field, body_field);
        const double force[3] = {fx, fy, fz};
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const IbmAdjShape sh =
                ibm_adj_shape<L, T>(n_nodes, n_markers, iterations, n_rules, n_tslots, slots_disjoint != 0);
            const Relaxation<T> r = make_relaxation<L, T>(omega_plus, omega_minus, force);
            const int pbx = std::min(256, 64 * ((nx + 63) / 64));
            const int pby = std::max(1, std::min(ny, sh.threads / pbx));
            const int tiles_x = (nx + pbx - 1) / pbx, tiles_y = (ny + pby - 1) / pby;
            const int64_t blocks = 1 + (int64_t)tiles_x * tiles_y * nz;
            require(blocks < (1ll << 31), "grid too large");
            dispatch_force_mask(force_mode, mask_mode, [&](auto fm, auto mk) {
                constexpr ForceMode FM = decltype(fm)::value;
                constexpr bool MK = decltype(mk)::value;
                auto go = [&](auto staged) {
                    constexpr bool ST = decltype(staged)::value;
                    if (ST) allow_lds(k_ibm_adj_pull<L, T, FM, MK, ST>, sh.lds);
                    k_ibm_adj_pull<L, T, FM, MK, ST><<<(unsigned)blocks, sh.threads, sh.lds, as_stream(stream)>>>(
                        LBM_IBM_ADJ_CALL_ARGS, static_cast<const T*>(pull_src), static_cast<T*>(pull_dst), kind,
                        static_cast<const T*>(wall_u), static_cast<const T*>(force_field),
                        static_cast<const T*>(body_field), g, r, pbx, pby, tiles_x, tiles_y);
                };
                if (sh.staged) go(std::true_type{});
                else go(std::false_type{});
            });
            check_launch("IBM adjoint with pull step");
        });
    });
}

namespace {

// The fused step's workgroup size bound and when it applies: float32 D2Q9, blocks of 64 ..
// 1024 lanes, and the prefetching IBM adjoint (slots disjoint, rows and rule adjoints staged).
bool adj_fused_ok(int lattice, int dtype, int n_nodes, int n_markers, int iterations, int n_tslots, int bx, int by)
{
    if (lattice != LBM_D2Q9 || dtype != LBM_F32 || bx <= 0 || by <= 0) return false;
    const int lanes = bx * by;
    if (lanes < RULE_CHUNK_MAX || lanes > 1024 || n_nodes <= 0 || n_markers <= 0 || iterations < 1) return false;
    const IbmAdjShape sh = ibm_adj_shape<D2Q9, float>(n_nodes, n_markers, iterations, 0, n_tslots, true);
    return sh.rows_in_lds > 0;
}

}  // namespace

LBM_API int lbm_adj_fused_supported(int lattice, int dtype, int n_nodes, int n_markers, int iterations, int n_tslots,
                                    int bx, int by, int* ok)
{
    return guarded("lbm_adj_fused_supported", [&] {
        require(ok != nullptr, "null output");
        *ok = adj_fused_ok(lattice, dtype, n_nodes, n_markers, iterations, n_tslots, bx, by) ? 1 : 0;
    });
}

LBM_API int lbm_adj_fused_step(
    int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff, const void* constant, int m,
    const int64_t* nodes, int n_nodes, const int* inverse, const void* weights, const void* areas, int n_markers,
    const int* spread_start, const int* spread_entry, int iterations, double fx, double fy, double fz, int64_t n_total,
    void* nu, const double* body_weight, void* lam, const int64_t* tslot, const int* tstart, const int* trow,
    const void* tcoeff, int n_tslots, double* grad_w, const void* record, const void* lam_next, const uint8_t* kind,
    const void* wall_u, const void* force_field, const void* body_field, int nx, int ny, int nz, double omega_plus,
    double omega_minus, double* grad_body, const int32_t* region, int n_region, int keep_pushes,
    const uint32_t* arule_info, const void* arule_coeff, int arule_nodes, int bx, int by, const void* pull_src,
    void* pull_dst, const void* pull_field, const uint32_t* rule_node_info, const int32_t* rule_node_first,
    int rule_nodes, const int64_t* prule_slot, const void* prule_coeff, const void* prule_constant, int64_t n_prules,
    int prule_m, const void* next_f, const void* next_record, const void* next_field, void* stream)
{
    return guarded("lbm_adj_fused_step", [&] {
        require(adj_fused_ok(lattice, dtype, n_nodes, n_markers, iterations, n_tslots, bx, by),
                "the fused adjoint step takes float32 D2Q9, 64 .. 1024 lanes and a body whose IBM adjoint stages "
                "its tables and transposed rows in LDS (lbm_adj_fused_supported)");
        check_ibm_adj_args(m, n_nodes, n_markers, iterations, n_total, 0,
                           {f, slot, coeff, constant, nodes, inverse, weights, areas, spread_start, spread_entry, nu,
                            body_weight, lam, tslot, tstart, trow, tcoeff, grad_w},
                           {});
        const Grid g = make_grid(lattice, nx, ny, nz);
        require((int64_t)g.n == n_total, "the grid must have n_total nodes");
        require(lam_next && kind && body_field && force_field && lam_next != lam, "null or aliased adjoint buffers");
        require(n_region >= 0 && (n_region == 0 || region), "null region list");
        require(arule_nodes >= 0 && (arule_nodes == 0 || (arule_info && arule_coeff)), "null link-rule node tables");
        require(reinterpret_cast<uintptr_t>(arule_info) % 16 == 0, "arule_info must be 16-byte aligned");
        require(!keep_pushes || (n_region == n_nodes && n_region <= bx * by &&
                                 (int64_t)n_tslots == (int64_t)n_region * (lattice == LBM_D2Q9 ? 9 : 19)),
                "keep_pushes needs the region to be the IBM node list, at most one node per lane, whose pushes are "
                "the transposed slots");
        const bool partner = pull_src != nullptr;
        if (partner) {
            require(pull_dst && pull_field && pull_src != pull_dst && pull_src != f && pull_dst != f,
                    "partner buffers must be distinct and non-null");
            require(n_prules >= 0 && n_prules < (1ll << 31) && rule_nodes >= 0 && (n_prules > 0) == (rule_nodes > 0),
                    "partner rules need rule nodes and nodes rows");
            require(n_prules == 0 || (rule_node_info && rule_node_first && prule_slot && prule_coeff &&
                                      prule_constant && prule_m >= 1 && prule_m <= 128),
                    "partner link rules need the node list, slot, coeff, constant and 1 <= m <= 128");
            require(reinterpret_cast<uintptr_t>(rule_node_info) % 16 == 0, "rule_node_info must be 16-byte aligned");
        }
        const double force[3] = {fx, fy, fz};
        const dim3 block(bx, by, 1);
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            if constexpr (std::is_same<T, float>::value && L::D == 2) {
                constexpr ForceMode FM = ForceMode::DENSE;
                const IbmAdjShape sh = ibm_adj_shape<L, T>(n_nodes, n_markers, iterations, 0, n_tslots, true);
                const Relaxation<T> r = make_relaxation<L, T>(omega_plus, omega_minus, force);
                const dim3 tiles = grid_of(g, block);
                const int tx = (int)tiles.x, ty = (int)tiles.y;
                const int64_t n_tiles = (int64_t)tiles.x * tiles.y * tiles.z;
                const int hosts = (int)std::min<int64_t>(n_tiles, 1 << 20);
                int chunk = RULE_CHUNK_MIN;
                int chunks = 0;
                if (partner && rule_nodes > 0) {
                    chunk = rule_chunk_nodes(rule_nodes, hosts);
                    chunks = rule_chunk_count(rule_nodes, chunk);
                }
                const RuleRows<T> rr{reinterpret_cast<const uint4*>(rule_node_info), rule_node_first,
                                     partner ? rule_nodes : 0, chunk, prule_slot, static_cast<const T*>(prule_coeff),
                                     static_cast<const T*>(prule_constant), partner ? n_prules : 0, prule_m};
                const int achunk = arule_nodes > 0 ? rule_chunk_nodes(arule_nodes, hosts) : RULE_CHUNK_MIN;
                const int achunks = arule_nodes > 0 ? rule_chunk_count(arule_nodes, achunk) : 0;
                const AdjRuleNodes<T> ar{reinterpret_cast<const uint4*>(arule_info),
                                         static_cast<const T*>(arule_coeff), arule_nodes, achunk};
                const int64_t ptiles = partner ? n_tiles : 0;
                const int64_t blocks = 1 + achunks + chunks + ptiles + n_tiles;
                require(blocks < (1ll << 31), "grid too large");
                // Kept pushes need no transposed rows in LDS.
                const int trows = keep_pushes ? 0 : sh.rows_in_lds;
                const size_t ibm_lds = adj_lds_plan(L::D, n_nodes, n_markers, iterations, sizeof(T), 0, trows).total;
                const size_t lds = std::max({ibm_lds, chunks ? sizeof(T) * (size_t)chunk * (L::Q - 1) : (size_t)0,
                                             achunks ? sizeof(T) * (size_t)achunk * L::Q * L::Q : (size_t)0});
                require(lds <= lds_limit(), "the workgroups' tables do not fit the LDS");
                // The next step's reads go through the last adjoint tile on the XCD of workgroup
                // 0 (workgroups are dealt to the 8 XCDs in turn), when all are given.
                AdjNext<T> next{static_cast<const T*>(next_f), static_cast<const T*>(next_record),
                                static_cast<const T*>(next_field), -1};
                if (next_f && next_record && next_field) {
                    int64_t last = blocks - 1;
                    last -= last % 8;
                    if (last >= blocks - n_tiles) next.block = (int)last;
                }
                allow_lds(k_adj_fused<L, T, FM>, lds);
                k_adj_fused<L, T, FM><<<(unsigned)blocks, block, lds, as_stream(stream)>>>(
                    static_cast<const T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), m,
                    nodes, n_nodes, inverse, static_cast<const T*>(weights), static_cast<const T*>(areas), n_markers,
                    spread_start, spread_entry, iterations, T(0.5 * fx), T(0.5 * fy), T(0.5 * fz), (uint32_t)n_total,
                    static_cast<const T*>(nu), body_weight, static_cast<T*>(lam), tslot, tstart, trow,
                    static_cast<const T*>(tcoeff), n_tslots, grad_w, nullptr, nullptr, 0, nullptr, nullptr, nullptr,
                    nullptr, 0, nullptr, 0, static_cast<const T*>(record), trows, static_cast<const T*>(lam_next),
                    kind, static_cast<const T*>(wall_u), static_cast<const T*>(force_field),
                    static_cast<const T*>(body_field), grad_body, region, n_region, keep_pushes, ar, g, r, tx, ty,
                    static_cast<const T*>(pull_src), static_cast<T*>(pull_dst), static_cast<const T*>(pull_field), rr,
                    tx, ty, (int)ptiles, next);
            } else {
                require(false, "the fused adjoint step takes float32 D2Q9");
            }
            check_launch("fused adjoint step");
        });
    });
}

#undef LBM_IBM_ADJ_CALL_ARGS
#undef LBM_IBM_ADJ_CHECK

}  // extern "C"
