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

#include "lbm_common.h"

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
    const T ux = (jx + T(0.5) * Gx) / rho;
    const T uy = (jy + T(0.5) * Gy) / rho;
    const T uz = L::D == 3 ? (jz + T(0.5) * Gz) / rho : T(0);
    const T alpha = r.half_wp + r.half_wm;
    const T beta = r.half_wp - r.half_wm;
    const bool guo = (FM != ForceMode::UNIFORM) || r.has_force;

    T drho_bar = T(0), rho_bar = T(0), gsum = T(0);
    T ubx = T(0), uby = T(0), ubz = T(0);
    T fbx = T(0), fby = T(0), fbz = T(0);
    const T usq = T(1.5) * (ux * ux + uy * uy + uz * uz);
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        using Di = Dir<L, i>;
        const T nbar = -alpha * lam[i] - beta * lam[L::opp[i]];
        fbar[i] = lam[i] + nbar;
        const T feqbar = -nbar;
        const T w = T(Di::w);
        const T cu = T(Di::x) * ux + T(Di::y) * uy + T(Di::z) * uz;
        const T g = T(3) * cu + T(4.5) * cu * cu - usq;
        drho_bar += feqbar * w;
        rho_bar += feqbar * w * g;
        const T gbar = feqbar * w * rho;
        gsum += gbar;
        const T s = gbar * (T(3) + T(9) * cu);
        ubx += s * T(Di::x);
        uby += s * T(Di::y);
        ubz += s * T(Di::z);
        if (guo) {
            const T cF = T(Di::x) * Gx + T(Di::y) * Gy + T(Di::z) * Gz;
            const T lw = lam[i] * w;
            const T lp = lw * r.a_plus;
            ubx += lp * (T(9) * cF * T(Di::x) - T(3) * Gx);
            uby += lp * (T(9) * cF * T(Di::y) - T(3) * Gy);
            ubz += lp * (T(9) * cF * T(Di::z) - T(3) * Gz);
            const T am3 = T(3) * r.a_minus;
            fbx += lp * (T(9) * cu * T(Di::x) - T(3) * ux) + lw * am3 * T(Di::x);
            fby += lp * (T(9) * cu * T(Di::y) - T(3) * uy) + lw * am3 * T(Di::y);
            fbz += lp * (T(9) * cu * T(Di::z) - T(3) * uz) + lw * am3 * T(Di::z);
        }
    });
    ubx -= T(3) * ux * gsum;
    uby -= T(3) * uy * gsum;
    ubz -= T(3) * uz * gsum;
    if constexpr (L::D == 2) ubz = T(0);
    const T jbx = ubx / rho, jby = uby / rho, jbz = ubz / rho;
    fbx += ubx / (T(2) * rho);
    fby += uby / (T(2) * rho);
    fbz += ubz / (T(2) * rho);
    rho_bar -= (ubx * ux + uby * uy + ubz * uz) / rho;
    drho_bar += rho_bar;
    static_for<0, L::Q>([&](auto I) {
        using Di = Dir<L, I>;
        fbar[I] += drho_bar + T(Di::x) * jbx + T(Di::y) * jby + T(Di::z) * jbz;
    });
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

template <class L, class T, ForceMode FM, bool MASK>
__global__ __launch_bounds__(256) void k_adj_pull(const T* __restrict__ lam_next, T* __restrict__ lam_out,
                                                  const T* __restrict__ f, const uint8_t* __restrict__ kind,
                                                  const T* __restrict__ wall_u, const T* __restrict__ field,
                                                  const T* __restrict__ body, double* __restrict__ grad_body,
                                                  T* __restrict__ nu_forced, Grid g, Relaxation<T> r)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const uint32_t node = ((uint32_t)z * (uint32_t)g.ny + (uint32_t)y) * (uint32_t)g.nx + (uint32_t)x;
    const uint8_t k = MASK ? kind[node] : 0;
    if (MASK && (k & KIND_SOLID)) return;
    uint32_t blocked = 0;
    if (MASK && (k & KIND_NEAR_SOLID)) {
        static_for<1, L::Q>([&](auto I) {
            if (kind[nbr<L, I, -1>(x, y, z, g)] & KIND_SOLID) blocked |= 1u << I;
        });
    }
    T fs[L::Q], lam[L::Q];
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        const uint32_t up = nbr<L, i, -1>(x, y, z, g);
        if (MASK && ((blocked >> i) & 1u))
            fs[i] = f[(size_t)L::opp[i] * g.n + node] + wall_term<L, i>(wall_u, up, g.n);
        else
            fs[i] = f[(size_t)i * g.n + up];
        lam[i] = lam_next[(size_t)i * g.n + node];
    });
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
        if (MASK && ((blocked >> i) & 1u))
            lam_out[(size_t)L::opp[i] * g.n + node] = mu[i];
        else
            lam_out[(size_t)i * g.n + nbr<L, i, -1>(x, y, z, g)] = mu[i];
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

// Transposed link rules: a_r = lam[target_r]; lam[target_r] = 0; lam[slot] += sum coeff a_r
// over the (rule, coefficient) entries of each distinct term slot (host-built, fixed order).
template <class T>
__global__ __launch_bounds__(1024) void k_adj_rules(T* __restrict__ lam, const int64_t* __restrict__ target,
                                                    int n_rules, const int64_t* __restrict__ uslot,
                                                    const int* __restrict__ ustart, const int* __restrict__ urule,
                                                    const T* __restrict__ ucoeff, int n_unique, T* __restrict__ a)
{
    for (int r = threadIdx.x; r < n_rules; r += blockDim.x) a[r] = lam[target[r]];
    __syncthreads();
    for (int r = threadIdx.x; r < n_rules; r += blockDim.x) lam[target[r]] = T(0);
    __syncthreads();
    for (int u = threadIdx.x; u < n_unique; u += blockDim.x) {
        T s = T(0);
        for (int e = ustart[u]; e < ustart[u + 1]; ++e) s += ucoeff[e] * a[urule[e]];
        lam[uslot[u]] += s;
    }
}

// ------------------------------------------------------------------ IBM adjoint

constexpr int IBM_ADJ_THREADS = 1024;

template <class T>
struct IbmAdjScratch {
    T *moments, *rho, *ustar, *rho_m;     // (d+1) nn, nn, d nn, nm
    T *force;                              // (iters + 1) d nn
    T *u;                                  // iters d nn
    T *U, *Fm;                             // iters d nm each
    T *force_bar, *u_bar, *ustar_bar, *rho_bar, *rho_m_bar, *Fm_bar, *U_bar, *m_bar;
};

template <class T>
__host__ __device__ inline size_t ibm_adj_layout(T* base, int d, int nn, int nm, int iters, IbmAdjScratch<T>* s)
{
    size_t o = 0;
    auto take = [&](T** p, size_t count) {
        if (s) *p = base + o;
        o += count;
    };
    T* dummy = nullptr;
    take(s ? &s->moments : &dummy, (size_t)(d + 1) * nn);
    take(s ? &s->rho : &dummy, nn);
    take(s ? &s->ustar : &dummy, (size_t)d * nn);
    take(s ? &s->rho_m : &dummy, nm);
    take(s ? &s->force : &dummy, (size_t)(iters + 1) * d * nn);
    take(s ? &s->u : &dummy, (size_t)iters * d * nn);
    take(s ? &s->U : &dummy, (size_t)iters * d * nm);
    take(s ? &s->Fm : &dummy, (size_t)iters * d * nm);
    take(s ? &s->force_bar : &dummy, (size_t)d * nn);
    take(s ? &s->u_bar : &dummy, (size_t)d * nn);
    take(s ? &s->ustar_bar : &dummy, (size_t)d * nn);
    take(s ? &s->rho_bar : &dummy, nn);
    take(s ? &s->rho_m_bar : &dummy, nm);
    take(s ? &s->Fm_bar : &dummy, (size_t)d * nm);
    take(s ? &s->U_bar : &dummy, (size_t)d * nm);
    take(s ? &s->m_bar : &dummy, (size_t)(d + 1) * nn);
    return o;
}

// One workgroup: recompute the IBM iterations from the state, then run them backwards.
// Seeds: nu (dL/dh at the forced nodes, d x n_total, written by k_adj_pull) and
// body_weight[d] = dL/d(body force). Outputs: lam[slot] += (M^T m_bar)[slot] through
// the transposed moment rows (distinct slots with their (row, coeff) lists), and
// grad_w[l][s] += dL/dw_ls in float64.
template <class T, int D>
__global__ __launch_bounds__(IBM_ADJ_THREADS) void k_ibm_adj(
    const T* __restrict__ F, const int64_t* __restrict__ slot, const T* __restrict__ coeff,
    const T* __restrict__ constant, int m, const int64_t* __restrict__ nodes, int nn, const int* __restrict__ inverse,
    const T* __restrict__ weights, const T* __restrict__ areas, int nm, const int* __restrict__ spread_start,
    const int* __restrict__ spread_entry, int iters, T hfx, T hfy, T hfz, uint32_t n_total,
    const T* __restrict__ nu, const double* __restrict__ body_weight, T* __restrict__ lam,
    const int64_t* __restrict__ tslot, const int* __restrict__ tstart, const int* __restrict__ trow,
    const T* __restrict__ tcoeff, int n_tslots, double* __restrict__ grad_w, T* __restrict__ scratch,
    const int64_t* __restrict__ rtarget, int n_rules, const int64_t* __restrict__ ruslot,
    const int* __restrict__ rustart, const int* __restrict__ rurule, const T* __restrict__ rucoeff, int n_runique,
    T* __restrict__ rule_a)
{
    constexpr int S = D == 2 ? 9 : 27;
    IbmAdjScratch<T> s;
    ibm_adj_layout<T>(scratch, D, nn, nm, iters, &s);
    const int tid = threadIdx.x;
    // Transposed link rules first (they act on slots the IBM rows never read).
    if (n_rules > 0) {
        for (int r = tid; r < n_rules; r += IBM_ADJ_THREADS) rule_a[r] = lam[rtarget[r]];
        __syncthreads();
        for (int r = tid; r < n_rules; r += IBM_ADJ_THREADS) lam[rtarget[r]] = T(0);
        __syncthreads();
        for (int u = tid; u < n_runique; u += IBM_ADJ_THREADS) {
            T acc = T(0);
            for (int e = rustart[u]; e < rustart[u + 1]; ++e) acc += rucoeff[e] * rule_a[rurule[e]];
            lam[ruslot[u]] += acc;
        }
        __syncthreads();
    }
    const int rows = (D + 1) * nn;
    const T hf[3] = {hfx, hfy, hfz};
    // ---- forward recompute (same expressions as k_ibm; summation order is irrelevant here)
    for (int r = tid; r < rows; r += IBM_ADJ_THREADS) {
        T acc = constant[r];
        for (int p = 0; p < m; ++p) {
            const int64_t e = (int64_t)p * rows + r;
            const T a = coeff[e];
            if (a != T(0)) acc += a * F[slot[e]];
        }
        s.moments[r] = acc;
    }
    __syncthreads();
    for (int n = tid; n < nn; n += IBM_ADJ_THREADS) {
        const T rho = T(1) + s.moments[n];
        s.rho[n] = rho;
        for (int a = 0; a < D; ++a) {
            s.ustar[a * nn + n] = (s.moments[(1 + a) * nn + n] + hf[a]) / rho;
            s.force[a * nn + n] = T(0);
        }
    }
    __syncthreads();
    for (int l = tid; l < nm; l += IBM_ADJ_THREADS) {
        T acc = T(0);
        for (int p = 0; p < S; ++p) acc += weights[l * S + p] * s.rho[inverse[l * S + p]];
        s.rho_m[l] = acc;
    }
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        const T* fo = s.force + (size_t)it * D * nn;
        T* fn = s.force + (size_t)(it + 1) * D * nn;
        T* u = s.u + (size_t)it * D * nn;
        T* U = s.U + (size_t)it * D * nm;
        T* Fm = s.Fm + (size_t)it * D * nm;
        for (int n = tid; n < nn; n += IBM_ADJ_THREADS)
            for (int a = 0; a < D; ++a) u[a * nn + n] = s.ustar[a * nn + n] + fo[a * nn + n] / (T(2) * s.rho[n]);
        __syncthreads();
        for (int l = tid; l < nm; l += IBM_ADJ_THREADS)
            for (int a = 0; a < D; ++a) {
                T acc = T(0);
                for (int p = 0; p < S; ++p) acc += weights[l * S + p] * u[a * nn + inverse[l * S + p]];
                U[a * nm + l] = acc;
                Fm[a * nm + l] = (T(2) * s.rho_m[l]) * (T(0) - acc) * areas[l];
            }
        __syncthreads();
        for (int n = tid; n < nn; n += IBM_ADJ_THREADS)
            for (int a = 0; a < D; ++a) {
                T acc = T(0);
                for (int e = spread_start[n]; e < spread_start[n + 1]; ++e) {
                    const int k = spread_entry[e];
                    acc += weights[k] * Fm[a * nm + k / S];
                }
                fn[a * nn + n] = fo[a * nn + n] + acc;
            }
        __syncthreads();
    }
    // ---- reverse sweep
    const T total_bar[3] = {T(-body_weight[0]), T(-body_weight[1]), T(D == 3 ? -body_weight[2] : 0.0)};
    for (int n = tid; n < nn; n += IBM_ADJ_THREADS) {
        for (int a = 0; a < D; ++a) {
            s.force_bar[a * nn + n] = nu[(size_t)a * n_total + nodes[n]];
            s.ustar_bar[a * nn + n] = T(0);
        }
        s.rho_bar[n] = T(0);
    }
    for (int l = tid; l < nm; l += IBM_ADJ_THREADS) s.rho_m_bar[l] = T(0);
    __syncthreads();
    for (int it = iters - 1; it >= 0; --it) {
        const T* fo = s.force + (size_t)it * D * nn;
        const T* u = s.u + (size_t)it * D * nn;
        const T* U = s.U + (size_t)it * D * nm;
        const T* Fm = s.Fm + (size_t)it * D * nm;
        // Fm_bar from the spread and the total; weight gradient of the spread.
        for (int l = tid; l < nm; l += IBM_ADJ_THREADS) {
            T fb[3] = {total_bar[0], total_bar[1], total_bar[2]};
            for (int p = 0; p < S; ++p) {
                const int n = inverse[l * S + p];
                double gw = 0.0;
                for (int a = 0; a < D; ++a) {
                    const T fbar = s.force_bar[a * nn + n];
                    fb[a] += weights[l * S + p] * fbar;
                    gw += (double)(Fm[a * nm + l] * fbar);
                }
                grad_w[l * S + p] += gw;
            }
            T rmb = T(0);
            for (int a = 0; a < D; ++a) {
                s.Fm_bar[a * nm + l] = fb[a];
                rmb += fb[a] * T(2) * (T(0) - U[a * nm + l]) * areas[l];
                s.U_bar[a * nm + l] = -fb[a] * T(2) * s.rho_m[l] * areas[l];
            }
            s.rho_m_bar[l] += rmb;
            // weight gradient of the interpolation U = sum w u
            for (int p = 0; p < S; ++p) {
                double gw = 0.0;
                for (int a = 0; a < D; ++a) gw += (double)(s.U_bar[a * nm + l] * u[a * nn + inverse[l * S + p]]);
                grad_w[l * S + p] += gw;
            }
        }
        __syncthreads();
        // u_bar (gather over the spread lists), then into ustar_bar, force_bar, rho_bar.
        for (int n = tid; n < nn; n += IBM_ADJ_THREADS) {
            for (int a = 0; a < D; ++a) {
                T acc = T(0);
                for (int e = spread_start[n]; e < spread_start[n + 1]; ++e) {
                    const int k = spread_entry[e];
                    acc += weights[k] * s.U_bar[a * nm + k / S];
                }
                const T two_rho = T(2) * s.rho[n];
                s.ustar_bar[a * nn + n] += acc;
                s.force_bar[a * nn + n] += acc / two_rho;
                s.rho_bar[n] -= acc * fo[a * nn + n] / (two_rho * s.rho[n]);
            }
        }
        __syncthreads();
    }
    // rho_m = sum_s w rho(x_ls): weight gradient and rho_bar.
    for (int l = tid; l < nm; l += IBM_ADJ_THREADS)
        for (int p = 0; p < S; ++p) grad_w[l * S + p] += (double)(s.rho_m_bar[l] * s.rho[inverse[l * S + p]]);
    __syncthreads();
    for (int n = tid; n < nn; n += IBM_ADJ_THREADS) {
        T acc = T(0);
        for (int e = spread_start[n]; e < spread_start[n + 1]; ++e) {
            const int k = spread_entry[e];
            acc += weights[k] * s.rho_m_bar[k / S];
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
    for (int t = tid; t < n_tslots; t += IBM_ADJ_THREADS) {
        T acc = T(0);
        for (int e = tstart[t]; e < tstart[t + 1]; ++e) acc += tcoeff[e] * s.m_bar[trow[e]];
        lam[tslot[t]] += acc;
    }
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
        require(mask_mode == 0 || (kind && wall_u), "mask_mode 1 needs kind and wall velocity");
        require(force_mode >= 0 && force_mode <= 2, "unknown force mode");
        require(force_mode != 1 || (force_field && mask_mode == 1), "field force mode needs the field and kind");
        require(force_mode != 2 || body_field, "dense force mode needs a body field");
        const double force[3] = {fx, fy, fz};
        const dim3 block(bx, by, 1);
        const dim3 grid = grid_of(g, block);
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            const Relaxation<T> r = make_relaxation<L, T>(omega_plus, omega_minus, force);
            auto go = [&](auto fm, auto mk) {
                k_adj_pull<L, T, decltype(fm)::value, decltype(mk)::value><<<grid, block, 0, as_stream(stream)>>>(
                    static_cast<const T*>(lam_next), static_cast<T*>(lam_out), static_cast<const T*>(f), kind,
                    static_cast<const T*>(wall_u), static_cast<const T*>(force_field),
                    static_cast<const T*>(body_field), grad_body, static_cast<T*>(nu_forced), g, r);
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
            check_launch("adjoint step");
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

LBM_API int lbm_ibm_adjoint(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                            const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                            const void* weights, const void* areas, int n_markers, const int* spread_start,
                            const int* spread_entry, int iterations, double fx, double fy, double fz, int64_t n_total,
                            const void* nu, const double* body_weight, void* lam, const int64_t* tslot,
                            const int* tstart, const int* trow, const void* tcoeff, int n_tslots, double* grad_w,
                            void* scratch, const int64_t* rule_target, int n_rules, const int64_t* rule_uslot,
                            const int* rule_ustart, const int* rule_urule, const void* rule_ucoeff, int n_rule_unique,
                            void* rule_scratch, void* stream)
{
    return guarded("lbm_ibm_adjoint", [&] {
        require(n_rules == 0 || (rule_target && rule_uslot && rule_ustart && rule_urule && rule_ucoeff && rule_scratch),
                "null rule pointer");
        require(n_nodes > 0 && n_markers > 0 && iterations >= 1 && m >= 1, "bad IBM sizes");
        require(n_total > 0 && n_total < (1ll << 31), "node count out of range");
        require(f && slot && coeff && constant && nodes && inverse && weights && areas && spread_start &&
                    spread_entry && nu && body_weight && lam && tslot && tstart && trow && tcoeff && grad_w && scratch,
                "null pointer");
        dispatch(lattice, dtype, [&](auto Lt, auto Tt) {
            using L = decltype(Lt);
            using T = decltype(Tt);
            k_ibm_adj<T, L::D><<<1, IBM_ADJ_THREADS, 0, as_stream(stream)>>>(
                static_cast<const T*>(f), slot, static_cast<const T*>(coeff), static_cast<const T*>(constant), m,
                nodes, n_nodes, inverse, static_cast<const T*>(weights), static_cast<const T*>(areas), n_markers,
                spread_start, spread_entry, iterations, T(0.5 * fx), T(0.5 * fy), T(0.5 * fz), (uint32_t)n_total,
                static_cast<const T*>(nu), body_weight, static_cast<T*>(lam), tslot, tstart, trow,
                static_cast<const T*>(tcoeff), n_tslots, grad_w, static_cast<T*>(scratch), rule_target, n_rules,
                rule_uslot, rule_ustart, rule_urule, static_cast<const T*>(rule_ucoeff), n_rule_unique,
                static_cast<T*>(rule_scratch));
            check_launch("IBM adjoint");
        });
    });
}

}  // extern "C"
