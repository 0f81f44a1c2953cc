// Per-node operators of the fused step, written for compile-time lattices.
//
// Each operator keeps the reference's floating-point operation order
// (LBM-C/collision.cu, LBM-C/boundary_conditions/*) but drops the products
// by 0 and +-1 that its runtime tables spell out, which changes at most the
// sign of an exact zero.  With contraction off (the FP64 build) the results
// are the reference's bit for bit; the FP32 builds use the same code.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lattice.hpp"

namespace lbmc {

// Loads and stores through the global address space.  All lattice buffers
// are device (global) memory; kernels that read their buffer pointers from
// device memory (the batch) would otherwise access them as generic (flat)
// pointers.
template <class T>
__device__ __forceinline__ T ldg(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

template <class T>
__device__ __forceinline__ void stg(T *p, T v)
{
    *(__attribute__((address_space(1))) T *)p = v;
}

// 1/x: FP32 uses the hardware reciprocal (1 ulp); FP64 divides exactly
template <class C>
__device__ __forceinline__ C recip(C x)
{
    if constexpr (sizeof(C) == 4) return __builtin_amdgcn_rcpf(x);
    else return C(1) / x;
}

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F &&fn)
{
    if constexpr (B < E) {
        fn(std::integral_constant<int, B>{});
        sfor<B + 1, E>(fn);
    }
}

// sum_d e_d(i) u_d, in d order, as the reference's e_mul_u accumulates it
// (collision.cu:36-47) without the final factor
template <class L, int I, class C>
__device__ __forceinline__ C edot(const C *u)
{
    C v = C(0);
    bool first = true;
    sfor<0, L::D>([&](auto dd) {
        constexpr int d = decltype(dd)::value;
        constexpr int e = e_of<L>(d, I);
        if constexpr (e != 0) {
            if (first) v = (e > 0) ? u[d] : -u[d];
            else v = (e > 0) ? v + u[d] : v - u[d];
            first = false;
        }
    });
    return v;
}

// u_square, collision.cu:23-34
template <class L, class C>
__device__ __forceinline__ C u_square(const C *u)
{
    C v = u[0] * u[0];
    v = v + u[1] * u[1];
    if constexpr (L::D > 2) v = v + u[2] * u[2];
    return v * C(1.5);
}

// BGK equilibrium, collision.cu:53-59
template <class L, class C>
__device__ __forceinline__ void equilibrium(C rho, const C *u, C *feq)
{
    const C usq = u_square<L>(u);
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const C eu = edot<L, i>(u) * C(3.);
        feq[i] = rho * C(L::w[i]) * (C(1.0) + eu + (C(0.5) * eu * eu) - usq);
    });
}

// u <- u + F/(2 rho), collision.cu:74-78
template <class L, class C>
__device__ __forceinline__ void guo_shift(C rho, const C *F, C *u)
{
    const C h = C(1) / (C(2) * rho);
    sfor<0, L::D>([&](auto dd) {
        constexpr int d = decltype(dd)::value;
        u[d] = u[d] + h * F[d];
    });
}

// Guo source term, collision.cu:82-99; tau is the unmodified relaxation time
template <class L, class C>
__device__ __forceinline__ void guo_source(C tau, const C *u, const C *F, C *ft)
{
    const C om = C(1) - (C(1) / (C(2) * tau));
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const C wo = C(L::w[i]) * om;
        // the reference writes (e0 u0)+(e1 u1)[+(e2 u2)] here, without a 0 start
        C eu = C(0);
        bool first = true;
        sfor<0, L::D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            constexpr int e = e_of<L>(d, i);
            if constexpr (e != 0) {
                if (first) eu = (e > 0) ? u[d] : -u[d];
                else eu = (e > 0) ? eu + u[d] : eu - u[d];
                first = false;
            }
        });
        C acc = C(0);
        sfor<0, L::D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            constexpr int e = e_of<L>(d, i);
            C a = (C(e) - u[d]) * C(3);
            if constexpr (e != 0) a = a + C(e * 9) * eu;
            const C fc = wo * a;
            if constexpr (d == 0) acc = fc * F[0];
            else acc = acc + fc * F[d];
        });
        ft[i] = acc;
    });
}

// Smagorinsky relaxation time, collision.cu:445-463 with q_bar starting at 0
template <class L, class C>
__device__ __forceinline__ C smagorinsky_tau(C tau, C rho, C c, const C *f, const C *feq)
{
    constexpr int D = L::D;
    C P[3][3] = {};
    sfor<0, D>([&](auto aa) {
        constexpr int a = decltype(aa)::value;
        sfor<a, D>([&](auto bb) {
            constexpr int b = decltype(bb)::value;
            C s = C(0);
            bool first = true;
            sfor<0, L::Q>([&](auto qq) {
                constexpr int q = decltype(qq)::value;
                constexpr int p = e_of<L>(a, q) * e_of<L>(b, q);
                if constexpr (p != 0) {
                    const C dq = f[q] - feq[q];
                    if (first) s = (p > 0) ? dq : -dq;
                    else s = (p > 0) ? s + dq : s - dq;
                    first = false;
                }
            });
            P[a][b] = s;
            P[b][a] = s;
        });
    });
    C qh = C(0);
    sfor<0, D>([&](auto aa) {
        constexpr int a = decltype(aa)::value;
        sfor<0, D>([&](auto bb) {
            constexpr int b = decltype(bb)::value;
            qh = qh + sqrt(C(2) * P[a][b] * P[a][b]);
        });
    });
    // rho*POW4(1/sqrt(3)) expands to rho*1/s*1/s*1/s*1/s, left to right
    const C s3 = C(1.7320508075688772);
    const C den = rho / s3 / s3 / s3 / s3;
    return tau + C(0.5) * (sqrt((tau * tau) + (C(2) * C(1.4142135623730951) * (c * c) *
                                               (C(1) / den) * qh)) -
                           tau);
}

// ---- FP32-compute forms.  The FP32/FP16S/BF16S paths compare to the FP64
// oracle within a tolerance, so they may regroup the arithmetic; these forms
// are algebraically identical to the reference's and cheaper:
//   * equilibrium: opposite directions share the even part,
//     feq_i = rho w (1 - usq + (eu)^2/2) +- rho w eu;
//   * Smagorinsky: sqrt(2 x^2) = sqrt(2)|x| and the symmetric Pi_ab = Pi_ba,
//     so Qhat = sqrt(2) (sum_a |Pi_aa| + 2 sum_{a<b} |Pi_ab|), and
//     2 sqrt(2) C^2 (9/rho) Qhat = 36 C^2 S / rho with S the bracket above;
//     tau + (sqrt(tau^2 + A) - tau)/2 = (tau + sqrt(tau^2 + A))/2.

template <class L, class C>
__device__ __forceinline__ void equilibrium_paired(C rho, const C *u, C *feq)
{
    const C base = C(1) - u_square<L>(u);
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int o = L::opp[i];
        if constexpr (i == 0) {
            feq[0] = rho * C(L::w[0]) * base;
        } else if constexpr (i < o) {
            const C eu = edot<L, i>(u) * C(3);
            const C rw = rho * C(L::w[i]);
            const C even = rw * (base + C(0.5) * eu * eu);
            const C odd = rw * eu;
            feq[i] = even + odd;
            feq[o] = even - odd;
        }
    });
}

template <class L, class C>
__device__ __forceinline__ C smagorinsky_tau_fast(C tau, C inv_rho, C c, const C *f, const C *feq)
{
    constexpr int D = L::D;
    C d[L::Q];
    sfor<0, L::Q>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        d[q] = f[q] - feq[q];
    });
    C sdiag = C(0), soff = C(0);
    sfor<0, D>([&](auto aa) {
        constexpr int a = decltype(aa)::value;
        sfor<a, D>([&](auto bb) {
            constexpr int b = decltype(bb)::value;
            C s = C(0);
            sfor<0, L::Q>([&](auto qq) {
                constexpr int q = decltype(qq)::value;
                constexpr int p = e_of<L>(a, q) * e_of<L>(b, q);
                if constexpr (p > 0) s += d[q];
                else if constexpr (p < 0) s -= d[q];
            });
            if constexpr (a == b) sdiag += fabs(s);
            else soff += fabs(s);
        });
    });
    const C A = C(36) * (c * c) * (sdiag + C(2) * soff) * inv_rho;
    return C(0.5) * (tau + sqrt(tau * tau + A));
}

// moment-space equilibria, collision.cu:465-503
template <class L, class C>
__device__ __forceinline__ void moment_equilibrium(C rho, const C *u, C qc, C *meq)
{
    if constexpr (L::D == 2) {
        const C jx = rho * u[0];
        const C jy = rho * u[1];
        meq[0] = rho;
        meq[1] = (C(-2) * rho) + C(3) * (jx * jx + jy * jy);
        meq[2] = rho - C(3) * (jx * jx + jy * jy);
        meq[3] = jx;
        meq[4] = -jx;
        meq[5] = jy;
        meq[6] = -jy;
        meq[7] = jx * jx - jy * jy;
        meq[8] = jx * jy;
    } else {
        const C jx = rho * u[0];
        const C jy = rho * u[1];
        const C jz = rho * u[2];
        const C j2 = jx * jx + jy * jy + jz * jz;
        meq[0] = rho;
        meq[1] = (-rho) + j2;
        meq[2] = rho - C(5) * j2;
        meq[3] = jx;
        meq[4] = qc * jx;
        meq[5] = jy;
        meq[6] = qc * jy;
        meq[7] = jz;
        meq[8] = qc * jz;
        meq[9] = (C(2) * jx * jx - (jy * jy + jz * jz));
        meq[10] = jy * jy - jz * jz;
        meq[11] = jx * jy;
        meq[12] = jy * jz;
        meq[13] = jx * jz;
        meq[14] = C(0);
    }
}

// Shifted storage (g = f - w): M g - (meq - M w) equals M f - meq, and
// M w is the equilibrium's moment vector at rest with rho = 1, so meq - M w
// is meq with rho replaced by s = rho - 1 in its rho-linear terms.  This
// keeps the MRT update on g free of the rounding of rho and of f = g + w.
template <class L>
constexpr bool mw_is_rest_equilibrium()
{
    // M w must be (1, -2, 1, 0, ...) for D2Q9 and (1, -1, 1, 0, ...) for D3Q15
    for (int k = 0; k < L::Q; k++) {
        double acc = 0;
        for (int j = 0; j < L::Q; j++) acc += L::M[k][j] * L::w[j];
        double want = 0;
        if (k == 0 || k == 2) want = 1;
        if (k == 1) want = L::D == 2 ? -2 : -1;
        const double d = acc - want;
        if (d > 1e-12 || d < -1e-12) return false;
    }
    return true;
}
static_assert(mw_is_rest_equilibrium<D2Q9>(), "D2Q9 M w");
static_assert(mw_is_rest_equilibrium<D3Q15>(), "D3Q15 M w");

template <class L, class C>
__device__ __forceinline__ void moment_equilibrium_shifted(C s, C rho, const C *u, C qc, C *meq)
{
    moment_equilibrium<L>(rho, u, qc, meq);
    if constexpr (L::D == 2) {
        const C jx = rho * u[0];
        const C jy = rho * u[1];
        const C j2 = jx * jx + jy * jy;
        meq[0] = s;
        meq[1] = (C(-2) * s) + C(3) * j2;
        meq[2] = s - C(3) * j2;
    } else {
        const C jx = rho * u[0];
        const C jy = rho * u[1];
        const C jz = rho * u[2];
        const C j2 = jx * jx + jy * jy + jz * jz;
        meq[0] = s;
        meq[1] = (-s) + j2;
        meq[2] = s - C(5) * j2;
    }
}

// m_k = s_k (sum_j M_kj f_j - meq_k), collision.cu:204-216
template <class L, class C>
__device__ __forceinline__ void mrt_moments(const C *f, const C *meq, const C *s, C *m)
{
    sfor<0, L::Q>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        C acc = C(0);
        bool first = true;
        sfor<0, L::Q>([&](auto jj) {
            constexpr int j = decltype(jj)::value;
            constexpr int c = L::M[k][j];
            if constexpr (c != 0) {
                C t;
                if constexpr (c == 1) t = f[j];
                else if constexpr (c == -1) t = -f[j];
                else t = C(c) * f[j];
                acc = first ? t : acc + t;
                first = false;
            }
        });
        m[k] = s[k] * (acc - meq[k]);
    });
}

// sum_k Minv_ik m_k, collision.cu:218-228
template <class L, bool MFIX, int I, class C>
__device__ __forceinline__ C minv_row(const C *m)
{
    C acc = C(0);
    bool first = true;
    sfor<0, L::Q>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr double c = MinvOf<L, MFIX>::at(I, k);
        if constexpr (c != 0.0) {
            const C t = C(c) * m[k];
            acc = first ? t : acc + t;
            first = false;
        }
    });
    return acc;
}

// ---------------------------------------------------------------- micro BCs

// zh_pressure_x / zh_pressure_X, d2q9_zh_defs.cu:4-27
template <bool EAST, class C>
__device__ __forceinline__ void zh_d2q9(C *f, C rho, C *u)
{
    if constexpr (!EAST) {
        u[0] = C(1.0) - ((C(1.0) / rho) * (f[0] + f[2] + f[4] + C(2.0) * (f[3] + f[6] + f[7])));
        u[1] = C(0.0);
        f[1] = f[3] + (C(2.0 / 3.0) * rho * u[0]);
        f[5] = f[7] - (C(0.5) * (f[2] - f[4])) + (C(1.0 / 6.0) * rho * u[0]);
        f[8] = f[6] + (C(0.5) * (f[2] - f[4])) + (C(1.0 / 6.0) * rho * u[0]);
    } else {
        u[0] = C(-1.0) + ((C(1.0) / rho) * (f[0] + f[2] + f[4] + C(2.0) * (f[1] + f[5] + f[8])));
        u[1] = C(0.0);
        f[3] = f[1] - (C(2.0 / 3.0) * rho * u[0]);
        f[6] = f[8] - (C(0.5) * (f[2] - f[4])) - (C(1.0 / 6.0) * rho * u[0]);
        f[7] = f[5] + (C(0.5) * (f[2] - f[4])) - (C(1.0 / 6.0) * rho * u[0]);
    }
}

// D3Q15 Zou-He faces as d3q15_zh_defs.cu:4-92 write them: normal axis, side,
// the velocity component the u-terms use (u[0] on face Z, :87-91), the
// summation order of the in-plane and outgoing sums, the axis unknown and its
// partner, the transverse differences T1 = f[ta]-f[tb], T2 = f[tc]-f[td], and
// four diagonals f_k = (f_p +- (1/12) rho u_v) - 1/4 (c1 T1 + c2 T2).
struct ZhFace {
    int axis, side, v;
    int s0[5], s1[5];
    int kn, pn, ta, tb, tc, td;
    int dg[4][4];
};
struct ZhRefTable {
    static constexpr ZhFace face[6] = {
    {0, 1, 0, {0, 3, 4, 5, 6}, {2, 14, 10, 12, 8}, 1, 2, 3, 4, 5, 6,
     {{7, 14, -1, 1}, {11, 10, -1, -1}, {9, 12, 1, 1}, {13, 8, 1, -1}}},
    {0, -1, 0, {0, 3, 4, 5, 6}, {1, 7, 11, 9, 13}, 2, 1, 3, 4, 5, 6,
     {{14, 7, 1, -1}, {10, 11, 1, 1}, {12, 9, -1, 1}, {8, 13, -1, -1}}},
    {1, 1, 1, {0, 1, 2, 5, 6}, {4, 14, 10, 9, 13}, 3, 4, 1, 2, 5, 6,
     {{7, 14, -1, 1}, {11, 10, -1, -1}, {12, 9, 1, -1}, {8, 13, 1, 1}}},
    {1, -1, 1, {0, 1, 2, 5, 6}, {3, 7, 11, 12, 8}, 4, 3, 1, 2, 5, 6,
     {{14, 7, 1, -1}, {10, 11, 1, 1}, {9, 12, -1, 1}, {13, 8, -1, -1}}},
    {2, 1, 2, {0, 1, 2, 3, 4}, {6, 14, 11, 12, 13}, 5, 6, 1, 2, 3, 4,
     {{7, 14, -1, 1}, {10, 11, 1, -1}, {9, 12, -1, -1}, {8, 13, 1, 1}}},
    {2, -1, 0, {0, 1, 2, 3, 4}, {5, 7, 10, 9, 8}, 6, 5, 1, 2, 3, 4,
     {{14, 7, 1, -1}, {11, 10, -1, 1}, {12, 9, 1, 1}, {13, 8, -1, -1}}},
    };
};

template <int FACE, class C>
__device__ __forceinline__ void zh_d3q15_ref(C *f, C rho, C *u)
{
    constexpr ZhFace z = ZhRefTable::face[FACE];
    const C s0 = f[z.s0[0]] + f[z.s0[1]] + f[z.s0[2]] + f[z.s0[3]] + f[z.s0[4]];
    const C s1 = f[z.s1[0]] + f[z.s1[1]] + f[z.s1[2]] + f[z.s1[3]] + f[z.s1[4]];
    C un;
    if constexpr (z.side > 0) un = C(1.0) - ((C(1.0) / rho) * (s0 + C(2.0) * s1));
    else un = C(-1.0) + ((C(1.0) / rho) * (s0 + C(2.0) * s1));
    u[0] = C(0.0);
    u[1] = C(0.0);
    u[2] = C(0.0);
    u[z.axis] = un;
    const C uv = u[z.v];
    const C T1 = f[z.ta] - f[z.tb];
    const C T2 = f[z.tc] - f[z.td];
    if constexpr (z.side > 0) f[z.kn] = f[z.pn] + (C(2.0 / 3.0) * rho * uv);
    else f[z.kn] = f[z.pn] - (C(2.0 / 3.0) * rho * uv);
    const C a = (C(1.0 / 12.0) * rho * uv);
    sfor<0, 4>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr int kd = z.dg[k][0], kp = z.dg[k][1], c1 = z.dg[k][2], c2 = z.dg[k][3];
        const C t = (c1 > 0 ? T1 : -T1) + (c2 > 0 ? T2 : -T2);
        if constexpr (z.side > 0) f[kd] = f[kp] + a - (C(0.25) * t);
        else f[kd] = f[kp] - a - (C(0.25) * t);
    });
}

// Corrected D3Q15 Zou-He on the face (AXIS, SIDE): closes mass and all three
// momentum components (the transverse ones vanish).
template <int AXIS, int SIDE, class C>
__device__ __forceinline__ void zh_d3q15_fixed(C *f, C rho, C *u)
{
    using L = D3Q15;
    C s0 = C(0), s1 = C(0);
    C T[3] = {C(0), C(0), C(0)};
    sfor<0, 15>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int ea = e_of<L>(AXIS, i);
        if constexpr (ea == 0) {
            s0 += f[i];
            sfor<0, 3>([&](auto dd) {
                constexpr int d = decltype(dd)::value;
                constexpr int e = e_of<L>(d, i);
                if constexpr (d != AXIS && e != 0) T[d] += C(e) * f[i];
            });
        } else if constexpr (ea == -SIDE) {
            s1 += f[i];
        }
    });
    const C un = SIDE > 0 ? C(1.0) - (s0 + C(2.0) * s1) / rho : C(-1.0) + (s0 + C(2.0) * s1) / rho;
    u[0] = C(0);
    u[1] = C(0);
    u[2] = C(0);
    u[AXIS] = un;
    sfor<0, 15>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        if constexpr (e_of<L>(AXIS, i) == SIDE) {
            constexpr int p = L::opp[i];
            constexpr bool diag = (AXIS != 0 && L::ex[i] != 0) || (AXIS != 1 && L::ey[i] != 0) ||
                                  (AXIS != 2 && L::ez[i] != 0);
            if constexpr (!diag) {
                f[i] = f[p] + C(SIDE) * C(2.0 / 3.0) * rho * un;
            } else {
                C t = C(0);
                sfor<0, 3>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    if constexpr (d != AXIS) t += C(e_of<L>(d, i)) * T[d];
                });
                f[i] = f[p] + C(SIDE) * C(1.0 / 12.0) * rho * un - C(0.25) * t;
            }
        }
    });
}

// Directions entering through face (axis, side): bit i set when
// e_axis(i) == side.
template <class L, int AXIS, int SIDE>
constexpr unsigned entering_mask()
{
    unsigned m = 0;
    for (int i = 0; i < L::Q; i++)
        if (e_of<L>(AXIS, i) == SIDE) m |= 1u << i;
    return m;
}

// Evaluate one micro BC on a copy of the node and keep the copy where the
// node's code matches.  Every face is evaluated (straight-line code, selects
// only): branching per code lets the compiler merge the branches' array
// stores through a computed address, which would move f[] to scratch memory
// for every node.  Micro-BC nodes are few, so the extra arithmetic is cheap.
template <class L, class C, class Fn>
__device__ __forceinline__ void micro_case(bool hit, C *f, C rho, C *u, Fn &&fn)
{
    C g[L::Q];
    C v[3] = {u[0], u[1], u[2]};
    sfor<0, L::Q>([&](auto ii) { g[decltype(ii)::value] = f[decltype(ii)::value]; });
    fn(g, rho, v);
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        f[i] = hit ? g[i] : f[i];
    });
    sfor<0, L::D>([&](auto dd) {
        constexpr int d = decltype(dd)::value;
        u[d] = hit ? v[d] : u[d];
    });
}

// micro_conditions[code-1], micro_bc.cuh:4-16; sf copies are one predicated
// load per entering direction.
template <class L, class C>
__device__ __forceinline__ void apply_micro(int code, C *f, C rho, C *u, const C *nb, bool zhfix)
{
    unsigned sfm = 0;
    if constexpr (L::D == 2) {
        micro_case<L>(code == 1, f, rho, u, [](C *g, C r, C *v) { zh_d2q9<false>(g, r, v); });
        micro_case<L>(code == 2, f, rho, u, [](C *g, C r, C *v) { zh_d2q9<true>(g, r, v); });
        sfm = code == 5 ? entering_mask<L, 0, 1>()
            : code == 6 ? entering_mask<L, 0, -1>()
            : code == 7 ? entering_mask<L, 1, 1>()
            : code == 8 ? entering_mask<L, 1, -1>() : 0u;
    } else {
        if (code <= 6) {
            if (!zhfix) {
                sfor<0, 6>([&](auto ff) {
                    constexpr int face = decltype(ff)::value;
                    micro_case<L>(code == face + 1, f, rho, u,
                                  [](C *g, C r, C *v) { zh_d3q15_ref<face>(g, r, v); });
                });
            } else {
                sfor<0, 6>([&](auto ff) {
                    constexpr int face = decltype(ff)::value;
                    micro_case<L>(code == face + 1, f, rho, u, [](C *g, C r, C *v) {
                        zh_d3q15_fixed<face / 2, (face % 2 == 0) ? 1 : -1>(g, r, v);
                    });
                });
            }
        }
        sfm = code == 7 ? entering_mask<L, 0, 1>()
            : code == 8 ? entering_mask<L, 0, -1>()
            : code == 9 ? entering_mask<L, 1, 1>()
            : code == 10 ? entering_mask<L, 1, -1>()
            : code == 11 ? entering_mask<L, 2, 1>()
            : code == 12 ? entering_mask<L, 2, -1>() : 0u;
    }
    if (sfm) {
        sfor<0, L::Q>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            if (sfm & (1u << i)) f[i] = nb[i];
        });
    }
}

}  // namespace lbmc
