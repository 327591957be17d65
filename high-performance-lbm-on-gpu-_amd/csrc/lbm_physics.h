// Node-local physics: shifted equilibrium, TRT/BGK collision, Guo source, Ladd wall term.
//
// Mathematics of the reference (src/lbm/equilibrium.py:40-58, collision.py:104-146,
// forcing.py:37-56, boundaries.py:297-322; CUDA twin kernels/common.cuh:41-120):
//
//   rho  = 1 + sum_i f~_i,  u = (sum_i c_i f_i + F/2) / rho
//   f~eq_i = w_i [drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u)]
//   f*_i = f_i - w+/2 (n_i + n_opp) - w-/2 (n_i - n_opp) + S_i,  n = f - feq
//   S_i  = w_i [(1 - w+/2)(9 (c.u)(c.F) - 3 u.F) + (1 - w-/2) 3 c.F]
//
// Every sum and product below is evaluated in the same order as the NumPy
// expressions of the reference, and the library is built with
// -ffp-contract=off, so each node update rounds exactly as the NumPy oracle
// does. Constants that NumPy forms in float64 from Python scalars (0.5 w,
// 1 - 0.5 w, c.F and the antisymmetric Guo term for a uniform force) are
// formed in double on the host and rounded once to real_t, as NumPy does.
#pragma once

#include "lbm_lattice.h"

namespace lbm {

constexpr int MAXQ = 19;

// Per-launch collision constants, built by make_relaxation() on the host.
template <class T>
struct Relaxation {
    T half_wp;        // T(0.5 * omega_plus)
    T half_wm;        // T(0.5 * omega_minus)
    T a_plus;         // T(1 - 0.5 * omega_plus)
    T a_minus;        // T(1 - 0.5 * omega_minus)
    T force[3];       // T(F0)
    T half_force[3];  // T(0.5 * F0)
    T cdotf[MAXQ];    // T(c_i . F0), c.F0 summed in double
    T anti[MAXQ];     // T((1 - 0.5 w-) * (3 c_i . F0)), all in double
    int has_force;    // uniform force not identically zero
};

// Force handling of a launch.
//   UNIFORM: F = F0 everywhere, Guo constants from the host (NumPy with scalar force).
//   FIELD:   F = F0 (+ field at FORCED nodes), all arithmetic in real_t (NumPy with
//            array force; the reference's IBM runs).
//   DENSE:   F = F0 + body[x] (+ field at FORCED nodes), real_t arithmetic.
enum class ForceMode : int { UNIFORM = 0, FIELD = 1, DENSE = 2 };

// sum_a c_a v_a over the non-zero components of direction I, in axis order;
// an all-zero direction gives +0.
template <class L, int I, class T>
__device__ __forceinline__ T cdot(T vx, T vy, T vz)
{
    using Di = Dir<L, I>;
    T s = T(0);
    bool first = true;
    auto add = [&](int c, T v) {
        if (c == 0) return;
        const T term = (c == 1) ? v : -v;
        s = first ? term : s + term;
        first = false;
    };
    add(Di::x, vx);
    add(Di::y, vy);
    if constexpr (L::D == 3) add(Di::z, vz);
    return s;
}

// Shifted equilibrium f~eq_I for given drho, rho, u and usq = 1.5 u.u.
template <class L, int I, class T>
__device__ __forceinline__ T shifted_feq(T drho, T rho, T ux, T uy, T uz, T usq)
{
    using Di = Dir<L, I>;
    T inner;
    if constexpr (Di::rest) {
        inner = T(0) - usq;
    } else {
        const T cu = T(3) * cdot<L, I>(ux, uy, uz);
        inner = (cu + (T(0.5) * cu) * cu) - usq;
    }
    return T(Di::w) * (drho + rho * inner);
}

// Density deviation and first moments of a population array, in the
// reference's order: drho = ((f0 + f1) + f2) + ..., j_a over non-zero c_a.
template <class L, class T>
__device__ __forceinline__ void moments(const T (&f)[L::Q], T& drho, T& jx, T& jy, T& jz)
{
    drho = f[0];
    static_for<1, L::Q>([&](auto I) { drho = drho + f[I]; });
    bool fx = true, fy = true, fz = true;
    jx = jy = jz = T(0);
    static_for<0, L::Q>([&](auto I) {
        using Di = Dir<L, I>;
        if constexpr (Di::x != 0) { const T t = Di::x == 1 ? f[I] : -f[I]; jx = fx ? t : jx + t; fx = false; }
        if constexpr (Di::y != 0) { const T t = Di::y == 1 ? f[I] : -f[I]; jy = fy ? t : jy + t; fy = false; }
        if constexpr (L::D == 3 && Di::z != 0) { const T t = Di::z == 1 ? f[I] : -f[I]; jz = fz ? t : jz + t; fz = false; }
    });
}

// One TRT + Guo collision of the Q values in f (in place).
// F is the node's force for FIELD/DENSE modes (ignored for UNIFORM).
template <class L, ForceMode FM, class T>
__device__ __forceinline__ void collide(T (&f)[L::Q], const Relaxation<T>& r, T Fx, T Fy, T Fz)
{
    T drho, jx, jy, jz;
    moments<L>(f, drho, jx, jy, jz);
    const T rho = T(1) + drho;
    T hx, hy, hz;
    if constexpr (FM == ForceMode::UNIFORM) {
        hx = r.half_force[0];
        hy = r.half_force[1];
        hz = r.half_force[2];
    } else {
        hx = T(0.5) * Fx;
        hy = T(0.5) * Fy;
        hz = T(0.5) * Fz;
    }
    const T ux = (jx + hx) / rho;
    const T uy = (jy + hy) / rho;
    const T uz = (L::D == 3) ? (jz + hz) / rho : T(0);
    T usum = ux * ux + uy * uy;
    if constexpr (L::D == 3) usum = usum + uz * uz;
    const T usq = T(1.5) * usum;

    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        constexpr int j = L::opp[i];
        if constexpr (j == i) {
            const T n = f[i] - shifted_feq<L, i>(drho, rho, ux, uy, uz, usq);
            f[i] = f[i] - r.half_wp * (n + n);
        } else if constexpr (i < j) {
            const T ni = f[i] - shifted_feq<L, i>(drho, rho, ux, uy, uz, usq);
            const T nj = f[j] - shifted_feq<L, j>(drho, rho, ux, uy, uz, usq);
            const T sym = r.half_wp * (ni + nj);
            const T asym = r.half_wm * (ni - nj);
            f[i] = (f[i] - sym) - asym;
            f[j] = (f[j] - sym) + asym;
        }
    });

    const bool guo = (FM != ForceMode::UNIFORM) || r.has_force;
    if (!guo) return;
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
    T udf = ux * Gx + uy * Gy;
    if constexpr (L::D == 3) udf = udf + uz * Gz;
    const T three_udf = T(3) * udf;
    static_for<0, L::Q>([&](auto I) {
        constexpr int i = I;
        using Di = Dir<L, i>;
        T sym, anti;
        if constexpr (Di::rest) {
            sym = T(0) - three_udf;
            anti = T(0);
        } else {
            const T cdu = cdot<L, i>(ux, uy, uz);
            T cdf;
            if constexpr (FM == ForceMode::UNIFORM) {
                cdf = r.cdotf[i];
                anti = r.anti[i];
            } else {
                cdf = cdot<L, i>(Gx, Gy, Gz);
                anti = r.a_minus * (T(3) * cdf);
            }
            sym = ((T(9) * cdu) * cdf) - three_udf;
        }
        f[i] = f[i] + T(Di::w) * (r.a_plus * sym + anti);
    });
}

// Ladd moving-wall term 2 w_i / cs^2 (c_i . u_w) with u_w read at solid node s.
template <class L, int I, class T>
__device__ __forceinline__ T wall_term(const T* __restrict__ wall_u, uint32_t s, uint32_t n_nodes)
{
    using Di = Dir<L, I>;
    if constexpr (Di::rest) {
        return T(0);
    } else {
        T c_dot = T(0);
        bool first = true;
        auto add = [&](int c, int axis) {
            if (c == 0) return;
            const T v = wall_u[(size_t)axis * n_nodes + s];
            const T t = (c == 1) ? v : -v;
            c_dot = first ? t : c_dot + t;
            first = false;
        };
        add(Di::x, 0);
        add(Di::y, 1);
        if constexpr (L::D == 3) add(Di::z, 2);
        return T(Di::wall) * c_dot;
    }
}

// Host: derive the collision constants from float64 parameters, in the
// order NumPy evaluates the same Python-scalar expressions.
template <class L, class T>
Relaxation<T> make_relaxation(double omega_plus, double omega_minus, const double force[3])
{
    Relaxation<T> r{};
    r.half_wp = T(0.5 * omega_plus);
    r.half_wm = T(0.5 * omega_minus);
    const double a_plus = 1.0 - 0.5 * omega_plus;
    const double a_minus = 1.0 - 0.5 * omega_minus;
    r.a_plus = T(a_plus);
    r.a_minus = T(a_minus);
    r.has_force = 0;
    for (int a = 0; a < 3; ++a) {
        const double fa = a < L::D ? force[a] : 0.0;
        r.force[a] = T(fa);
        r.half_force[a] = T(0.5 * fa);
        if (fa != 0.0) r.has_force = 1;
    }
    for (int i = 0; i < MAXQ; ++i) {
        r.cdotf[i] = T(0);
        r.anti[i] = T(0);
    }
    for (int i = 0; i < L::Q; ++i) {
        const int c[3] = {L::cx[i], L::cy[i], L::cz[i]};
        double s = 0.0;
        bool first = true;
        for (int a = 0; a < L::D; ++a) {
            if (c[a] == 0) continue;
            const double t = c[a] * (a < L::D ? force[a] : 0.0);
            s = first ? t : s + t;
            first = false;
        }
        r.cdotf[i] = T(s);
        r.anti[i] = T(a_minus * (3.0 * s));
    }
    return r;
}

}  // namespace lbm
