// FP32-compute fluid-node update for the BGK family (bgk, bgk_guo and their
// Smagorinsky forms, LBM-C/collision.cu:49-111,445-463) with the arithmetic
// regrouped for throughput.  Used by the FP32/FP16S/BF16S kernels at plain
// fluid nodes (class 0, no partial saturation); boundary and porous nodes
// take the reference-order path.  Every regrouping below is an identity of
// the method, so results differ from the reference only by FP32 rounding:
//
//  * moments: rho, j and the second moment P_ab = sum e_a e_b f come from a
//    Walsh-Hadamard transform of the diagonal (corner) populations plus the
//    axis pairs (39 add/sub for D3Q15 instead of ~105);
//  * Smagorinsky: Pi = sum e e (f - feq) = P - rho (u u + I/3) because the
//    equilibrium's second moment is exactly rho (u u + I/3); the norm
//    sum_ab sqrt(2 Pi_ab^2) = sqrt(2) (sum |Pi_aa| + 2 sum_{a<b} |Pi_ab|);
//  * shifted storage (g = f - w): the update runs on g directly,
//    g' = g + omega (geq - g) with geq = feq - w, so no +-w conversions, and
//    geq is formed from s = rho - 1 (rho X - 1 = s X + (X - 1)) so that the
//    rounding of rho does not enter the non-equilibrium part;
//  * opposite directions share the even part of feq and of the Guo source.
#pragma once

#include "lattice.hpp"
#include "node_ops.hpp"
#include "step_args.hpp"

namespace lbmc {

// sum v, j = sum e v, P = sum e e v over the stored values v (g or f).
// P is packed xx, yy, zz, xy, yz, xz.
template <class L>
struct FastMoments;

template <>
struct FastMoments<D3Q15> {
    __device__ static __forceinline__ void run(const float *v, float &s, float *j, float *P)
    {
        // corners 7..14: (sx, sy, sz) = +++, -++, +-+, --+, ++-, -+-, +--, ---
        const float p1 = v[7] + v[8], m1 = v[7] - v[8];
        const float p2 = v[9] + v[10], m2 = v[9] - v[10];
        const float p3 = v[11] + v[12], m3 = v[11] - v[12];
        const float p4 = v[13] + v[14], m4 = v[13] - v[14];
        const float pp1 = p1 + p2, pm1 = p1 - p2, pp2 = p3 + p4, pm2 = p3 - p4;
        const float mp1 = m1 + m2, mm1 = m1 - m2, mp2 = m3 + m4, mm2 = m3 - m4;
        const float S0 = pp1 + pp2, Sz = pp1 - pp2;
        const float Sy = pm1 + pm2, Syz = pm1 - pm2;
        const float Sx = mp1 + mp2, Sxz = mp1 - mp2;
        const float Sxy = mm1 + mm2;
        const float axp = v[1] + v[2], axm = v[1] - v[2];
        const float ayp = v[3] + v[4], aym = v[3] - v[4];
        const float azp = v[5] + v[6], azm = v[5] - v[6];
        s = v[0] + (axp + ayp) + (azp + S0);
        j[0] = axm + Sx;
        j[1] = aym + Sy;
        j[2] = azm + Sz;
        P[0] = axp + S0;
        P[1] = ayp + S0;
        P[2] = azp + S0;
        P[3] = Sxy;
        P[4] = Syz;
        P[5] = Sxz;
    }
};

template <>
struct FastMoments<D2Q9> {
    __device__ static __forceinline__ void run(const float *v, float &s, float *j, float *P)
    {
        // corners 5..8: (sx, sy) = ++, -+, --, +-
        const float p1 = v[5] + v[6], m1 = v[5] - v[6];
        const float p2 = v[8] + v[7], m2 = v[8] - v[7];
        const float S0 = p1 + p2, Sy = p1 - p2, Sx = m1 + m2, Sxy = m1 - m2;
        const float axp = v[1] + v[3], axm = v[1] - v[3];
        const float ayp = v[2] + v[4], aym = v[2] - v[4];
        s = v[0] + (axp + ayp) + S0;
        j[0] = axm + Sx;
        j[1] = aym + Sy;
        j[2] = 0.f;
        P[0] = axp + S0;
        P[1] = ayp + S0;
        P[2] = 0.f;
        P[3] = Sxy;
        P[4] = 0.f;
        P[5] = 0.f;
    }
};

// Moments and velocity of a fluid node from its stored values v (f, or
// g = f - w when SH); u carries the Guo half-force shift (collision.cu:74-78).
template <class L>
struct FastState {
    float s;  // sum of the stored values: rho - 1 when shifted, rho otherwise
    float rho, ir;
    float u[3];
    float P[6];  // second moment of the stored values: xx, yy, zz, xy, yz, xz
};

template <class L, bool SH, bool GUO>
__device__ __forceinline__ void fast_prepare(const float *v, const float *F, FastState<L> &st)
{
    constexpr int D = L::D;
    constexpr float sh = SH ? 1.f : 0.f;
    float s, j[3];
    FastMoments<L>::run(v, s, j, st.P);
    st.s = s;
    st.rho = sh + s;
    st.ir = recip(st.rho);
    st.u[0] = j[0] * st.ir;
    st.u[1] = j[1] * st.ir;
    st.u[2] = D > 2 ? j[2] * st.ir : 0.f;
    if constexpr (GUO) {
        const float h = 0.5f * st.ir;
        sfor<0, D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            st.u[d] = st.u[d] + h * F[d];
        });
    }
}

// Smagorinsky relaxation time (collision.cu:445-463):
// Pi_ab = P_ab - rho u_a u_b - delta_ab (rho - sh)/3, norm sum_ab sqrt(2 Pi_ab^2)
template <class L, bool SH>
__device__ __forceinline__ float fast_smag_tau(const FastState<L> &st, float tau, float c_smag)
{
    const float rho = st.rho;
    const float *u = st.u;
    const float *P = st.P;
    const float dg = st.s * (1.f / 3.f);  // (rho - sh) / 3 without the rounding of rho
    float sd = fabsf(P[0] - rho * u[0] * u[0] - dg) + fabsf(P[1] - rho * u[1] * u[1] - dg);
    float so = fabsf(P[3] - rho * u[0] * u[1]);
    if constexpr (L::D > 2) {
        sd += fabsf(P[2] - rho * u[2] * u[2] - dg);
        so += fabsf(P[4] - rho * u[1] * u[2]) + fabsf(P[5] - rho * u[0] * u[2]);
    }
    const float A = 36.f * (c_smag * c_smag) * (sd + 2.f * so) * st.ir;
    return 0.5f * (tau + sqrtf(tau * tau + A));
}

// Non-equilibrium momentum flux per unit density, Pi_ab / rho (xx, yy, zz,
// xy, yz, xz; the z components are zero on D2Q9), and rho: the inputs of
// the learned relaxation-time closure.
template <class L, bool SH>
__device__ __forceinline__ void fast_closure_features(const FastState<L> &st, float (&x)[7])
{
    const float rho = st.rho;
    const float *u = st.u;
    const float *P = st.P;
    const float dg = st.s * (1.f / 3.f);
    x[0] = (P[0] - rho * u[0] * u[0] - dg) * st.ir;
    x[1] = (P[1] - rho * u[1] * u[1] - dg) * st.ir;
    x[3] = (P[3] - rho * u[0] * u[1]) * st.ir;
    if constexpr (L::D > 2) {
        x[2] = (P[2] - rho * u[2] * u[2] - dg) * st.ir;
        x[4] = (P[4] - rho * u[1] * u[2]) * st.ir;
        x[5] = (P[5] - rho * u[0] * u[2]) * st.ir;
    } else {
        x[2] = x[4] = x[5] = 0.f;
    }
    x[6] = rho;
}

// BGK relaxation with relaxation time t (and the Guo source with the base
// tau, collision.cu:69-111) on the stored values.
template <class L, bool SH, bool GUO>
__device__ __forceinline__ void fast_relax(float *v, const FastState<L> &st, float t, float tau,
                                           const float *F)
{
    constexpr int D = L::D;
    constexpr int Q = L::Q;
    constexpr float sh = SH ? 1.f : 0.f;
    const float rho = st.rho;
    const float *u = st.u;
    const float om = recip(t);
    const float usq = u_square<L>(u);
    const float base = 1.f - usq;
    const float u3[3] = {3.f * u[0], 3.f * u[1], 3.f * u[2]};
    // Guo: S_i = c w_i [3 (e_i - u).F + 9 (e_i.u)(e_i.F)], c = 1 - 1/(2 tau)
    float gc = 0.f, uF = 0.f;
    if constexpr (GUO) {
        gc = 1.f - 0.5f * recip(tau);
        uF = u[0] * F[0] + u[1] * F[1];
        if constexpr (D > 2) uF += u[2] * F[2];
    }
    // rest population; rho X - sh = s X + sh (X - 1) keeps the shifted
    // equilibrium free of the rounding of rho = 1 + s
    const float s = st.s;
    {
        const float veq = float(L::w[0]) * (s * base - sh * usq);
        float r = v[0] + om * (veq - v[0]);
        if constexpr (GUO) r += gc * float(L::w[0]) * (-3.f * uF);
        v[0] = r;
    }
    sfor<1, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int o = L::opp[i];
        if constexpr (i < o) {
            constexpr float w = float(L::w[i]);
            const float eu = edot<L, i>(u3);
            const float h = 0.5f * eu * eu;
            const float even = w * (s * (base + h) + sh * (h - usq));
            const float odd = w * rho * eu;
            float ri = v[i] + om * ((even + odd) - v[i]);
            float ro = v[o] + om * ((even - odd) - v[o]);
            if constexpr (GUO) {
                const float eF = edot<L, i>(F);
                const float se = gc * w * (3.f * eu * eF - 3.f * uF);
                const float so = gc * w * 3.f * eF;
                ri += se + so;
                ro += se - so;
            }
            v[i] = ri;
            v[o] = ro;
        }
    });
}

// Fluid-node BGK update on stored values v (f, or g = f - w when SH).
// F is the node's force (Guo); returns the stored macros in rho / u.
template <class L, bool SH, bool GUO, bool SMAG>
__device__ __forceinline__ void fluid_bgk_fast(float *v, float tau, float c_smag, const float *F,
                                               float &rho_out, float *u_out)
{
    FastState<L> st;
    fast_prepare<L, SH, GUO>(v, F, st);
    float t = tau;
    if constexpr (SMAG) {
        if (c_smag > 0.f) t = fast_smag_tau<L, SH>(st, tau, c_smag);
    }
    fast_relax<L, SH, GUO>(v, st, t, tau, F);
    rho_out = st.rho;
    sfor<0, L::D>([&](auto dd) { u_out[decltype(dd)::value] = st.u[decltype(dd)::value]; });
}

}  // namespace lbmc
