// The fused pull-stream / boundary / collide / store kernel.
//
// One launch performs one call of the reference's iterate_kernel
// (LBM-C/solver.cu:10-136) for every node:
//   1. index-shift gather: population i of node x is read from slot
//      p_i(x,t) = (x - t e_i) mod L in every axis (solver.cu:77-94);
//   2. rho = sum f_i, u = sum e_i f_i / rho (solver.cu:95-108);
//   3. node class from the 1-byte mask: solid -> bounceback; otherwise macro
//      BC, micro BC (solver.cu:111-112) and the collision (solver.cu:115);
//   4. the post-collision populations go back to the same slots
//      (solver.cu:119-123), so the update is in place with one buffer and
//      no ghost layers: each slot is owned by exactly one node per call;
//   5. rho and u are stored when asked (solver.cu:125-134).
//
// Work decomposition: one thread per node in slot order.  For a fixed
// direction a wave's 64 nodes read 64 consecutive slots of the source row
// (rotated by t*e_x), so the shift costs at most one extra cache line per
// row.  No __syncthreads and no LDS are needed: nodes are independent within
// a call.
#pragma once

#include <hip/hip_runtime.h>

#include "fast_bgk.hpp"
#include "lattice.hpp"
#include "node_index.hpp"
#include "node_ops.hpp"
#include "step_args.hpp"
#include "storage.hpp"

namespace lbmc {

// One node of iterate_kernel after the gather (solver.cu:95-134): v[] holds
// the node's pulled populations as stored (f_i, or g_i = f_i - w_i for the
// shifted formats) converted to the compute type; on return v[] holds the
// post-collision values in the same form, ready to be written back to the
// same slots.  Stores rho/u when the call asks for it.
template <class L, class St, int FAM, bool GUO, bool SMAG, bool MFIX>
__device__ __forceinline__ void node_update(const StepArgs<typename St::C> &a, int n,
                                            typename St::C (&v)[L::Q])
{
    using C = typename St::C;
    constexpr int Q = L::Q;
    constexpr int D = L::D;
    const C qc = (a.quirks & QK_FIX_D3Q15_MEQ) ? C(-7.0 / 3.0) : C(-2);
    const bool zhfix = (a.quirks & QK_FIX_D3Q15_ZH) != 0;

    // FP32 compute: plain fluid nodes of the BGK family take the regrouped
    // update (fast_bgk.hpp); everything else follows the reference's order.
    constexpr bool FAST = std::is_same<C, float>::value;
    const unsigned m = a.all_fluid ? 0u : ldg(a.mask + n);
    if constexpr (FAST && FAM == 0) {
        if (m == 0) {
            C F[3] = {C(0), C(0), C(0)};
            if constexpr (GUO) {
                sfor<0, D>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    F[d] = a.force[0] ? ldg(a.force[d] + n) : a.F[d];
                });
            }
            C rho, u[3];
            fluid_bgk_fast<L, St::shifted, GUO, SMAG>(v, a.tau, a.c_smag, F, rho, u);
            if (a.store) {
                sfor<0, D>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    stg(a.u[d] + n, u[d]);
                });
                stg(a.rho + n, rho);
            }
            return;
        }
    }

    // ---- moments, in the reference's summation order
    C f[Q];
    C rho = C(0);
    C j[3] = {C(0), C(0), C(0)};
    bool jfirst[3] = {true, true, true};
    sfor<0, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        rho = (i == 0) ? v[i] : rho + v[i];
        sfor<0, D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            constexpr int e = e_of<L>(d, i);
            if constexpr (e != 0) {
                if (jfirst[d]) j[d] = (e > 0) ? v[i] : -v[i];
                else j[d] = (e > 0) ? j[d] + v[i] : j[d] - v[i];
                jfirst[d] = false;
            }
        });
        if constexpr (St::shifted) f[i] = v[i] + C(L::w[i]);
        else f[i] = v[i];
    });
    C srho = rho;  // rho - 1 for shifted storage, exactly as summed
    if constexpr (St::shifted) rho = C(1) + rho;
    // FP32 compute regroups the arithmetic (one reciprocal of rho, paired
    // equilibria, the closed-form Smagorinsky norm); FP64 keeps the
    // reference's operation order bit for bit.
    C u[3];
    if constexpr (FAST) {
        const C ir = recip(rho);
        sfor<0, D>([&](auto dd) { u[decltype(dd)::value] = j[decltype(dd)::value] * ir; });
    } else {
        sfor<0, D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            u[d] = j[d] / rho;
        });
    }

    // ---- node class
    const uint16_t desc = m ? ldg(a.cls + m) : (uint16_t)0;

    bool stored_done = false;
    if (desc & CLS_SOLID) {
        // bounceback, collision.cu:423-443
        if constexpr (St::shifted) {
            // w_opp(i) = w_i: permute the stored values themselves (exact)
            C t[Q];
            sfor<0, Q>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                t[i] = v[L::opp[i]];
            });
            sfor<0, Q>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                v[i] = t[i];
            });
            stored_done = true;
        } else {
            C t[Q];
            sfor<0, Q>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                t[i] = f[L::opp[i]];
            });
            sfor<0, Q>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                f[i] = t[i];
            });
        }
        rho = C(0);
        sfor<0, D>([&](auto dd) { u[decltype(dd)::value] = C(0); });
    } else {
        C B = C(0);
        if (desc) {
            if (desc & CLS_POROUS) B = ldg(a.porosity + n);
            const int mc = cls_macro(desc);
            if (mc) {
                // presc_*, macro_bc.cu:19-125
                const unsigned ps = macro_set(mc);
                if (ps & 1u) {
                    rho = ldg(a.rho + n);
                    srho = St::shifted ? rho - C(1) : rho;
                }
                sfor<0, D>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    if (ps & (2u << d)) u[d] = ldg(a.u[d] + n);
                });
            }
            const int mi = cls_micro(desc);
            if (mi) {
                const C *nb = a.sfidx ? a.sfbuf + (size_t)a.sfidx[n] * Q : nullptr;
                apply_micro<L>(mi, f, rho, u, nb, zhfix);
            }
        }

        C F[3] = {C(0), C(0), C(0)};
        if constexpr (GUO) {
            if (a.force[0]) {
                sfor<0, D>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    F[d] = ldg(a.force[d] + n);
                });
            } else {
                sfor<0, D>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    F[d] = a.F[d];
                });
            }
        }

        if constexpr (FAM == 0) {
            // bgk / bgk_guo / bgk_ntpor / bgk_ntpor_guo, collision.cu:49-190
            C ft[Q];
            C tau = a.tau;
            if constexpr (GUO) {
                guo_shift<L>(rho, F, u);
                guo_source<L>(tau, u, F, ft);
            }
            C feq[Q];
            if constexpr (FAST) equilibrium_paired<L>(rho, u, feq);
            else equilibrium<L>(rho, u, feq);
            if constexpr (SMAG) {
                if (a.c_smag > C(0)) {
                    if constexpr (FAST) tau = smagorinsky_tau_fast<L>(tau, recip(rho), a.c_smag, f, feq);
                    else tau = smagorinsky_tau<L>(tau, rho, a.c_smag, f, feq);
                }
            }
            const C omega = FAST ? recip(tau) : C(1.0) / tau;
            if (a.porous) {
                const C ob = C(1) - B;
                C t[Q];
                sfor<0, Q>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    const C cb = omega * (f[i] - feq[i]);
                    const C cs = f[L::opp[i]] - f[i];
                    C r = f[i] - ob * cb + B * cs;
                    if constexpr (GUO) r = r + ob * ft[i];
                    t[i] = r;
                });
                sfor<0, Q>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    f[i] = t[i];
                });
            } else {
                sfor<0, Q>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    C r;
                    if constexpr (FAST) r = f[i] + omega * (feq[i] - f[i]);
                    else r = f[i] - omega * (f[i] - feq[i]);
                    if constexpr (GUO) r = r + ft[i];
                    f[i] = r;
                });
            }
        } else {
            // mrt / mrt_guo / mrt_ntpor / mrt_ntpor_guo, collision.cu:192-421
            C ft[Q];
            if constexpr (GUO) {
                guo_shift<L>(rho, F, u);
                guo_source<L>(a.tau, u, F, ft);
            }
            C meq[Q], mm[Q];
            bool gspace = false;
            if constexpr (St::shifted) gspace = cls_micro(desc) == 0;
            if (gspace) {
                // on the stored g = f - w (no micro BC changed f): the same
                // update, M g - (meq - M w) = M f - meq, and f[opp] - f =
                // g[opp] - g because w_opp(i) = w_i
                moment_equilibrium_shifted<L>(srho, rho, u, qc, meq);
                mrt_moments<L>(v, meq, a.tau_mrt, mm);
                if (a.porous) {
                    const C ob = C(1) - B;
                    sfor<0, Q>([&](auto ii) {
                        constexpr int i = decltype(ii)::value;
                        const C acc = minv_row<L, MFIX, i>(mm);
                        const C cs = v[L::opp[i]] - v[i];
                        C r = v[i] - ob * acc;
                        if constexpr (GUO) r = r + ob * ft[i];
                        v[i] = r + B * cs;
                    });
                } else {
                    sfor<0, Q>([&](auto ii) {
                        constexpr int i = decltype(ii)::value;
                        C r = v[i] - minv_row<L, MFIX, i>(mm);
                        if constexpr (GUO) r = r + ft[i];
                        v[i] = r;
                    });
                }
                stored_done = true;
            } else {
            moment_equilibrium<L>(rho, u, qc, meq);
            mrt_moments<L>(f, meq, a.tau_mrt, mm);
            if (a.porous) {
                // in place: f[opp[i]] is already post-collision when
                // opp[i] < i (collision.cu:331-343, :406-419)
                const C ob = C(1) - B;
                sfor<0, Q>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    const C acc = minv_row<L, MFIX, i>(mm);
                    const C cs = f[L::opp[i]] - f[i];
                    C r = f[i] - ob * acc;
                    if constexpr (GUO) r = r + ob * ft[i];
                    f[i] = r + B * cs;
                });
            } else {
                sfor<0, Q>([&](auto ii) {
                    constexpr int i = decltype(ii)::value;
                    C r = f[i] - minv_row<L, MFIX, i>(mm);
                    if constexpr (GUO) r = r + ft[i];
                    f[i] = r;
                });
            }
            }
        }
    }

    if (!stored_done) {
        sfor<0, Q>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            if constexpr (St::shifted) v[i] = f[i] - C(L::w[i]);
            else v[i] = f[i];
        });
    }
    if (a.store) {
        sfor<0, D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            stg(a.u[d] + n, u[d]);
        });
        stg(a.rho + n, rho);
    }
}

// n / d for 0 <= n < 2^31 with the magic pair (m, l) of d
__device__ __forceinline__ int fast_div(int n, unsigned m, int l)
{
    return (int)((__umulhi((unsigned)n, m) + (unsigned)n) >> l);
}

// 64-bit-offset kernels on lattices with Lx % 64 == 0: the 64 nodes of a
// wave lie in one x-row, so for direction i the start of the slot row,
// (n - x) + dd[1][e_y+1] + dd[2][e_z+1], is the same for the whole wave and
// goes into scalar registers; each lane keeps only its position in the row,
// x + dd[0][e_x+1] < Lx, a 32-bit byte offset.  This is the uniform-base
// plus 32-bit-offset access of the narrow kernels: on a 1024^3 FP32 lattice
// the step runs 25-30 % faster than with per-lane 64-bit addresses.
template <class L, class St, class F>
__device__ __forceinline__ void wide_row_slots(const StepArgs<typename St::C> &a, int n, const int (&dd)[3][3], F &&f)
{
    using S = typename St::S;
    const int x = n - fast_div(n, a.mx, a.lx) * a.L[0];
    const int r0 = n - x;
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const int rs = __builtin_amdgcn_readfirstlane(r0 + dd[1][L::ey[i] + 1] + dd[2][L::ez[i] + 1]);
        char *base = (char *)((S *)a.f + (size_t)i * a.npad + (size_t)rs);
        f(ii, (S *)(base + (unsigned)(x + dd[0][L::ex[i] + 1]) * (unsigned)sizeof(S)));
    });
}

// A lane's 32-bit byte offset, opaque to the compiler: its widening to 64
// bits then happens beside the access that uses it, so a store in another
// basic block than the gather that computed the same offset still takes a
// scalar base and the 32-bit offset (otherwise the compiler carries the
// widened offset over and adds 64-bit per-lane addresses: two vector
// instructions and two registers per direction).  No instruction.
__device__ __forceinline__ unsigned lane_offset(unsigned off)
{
    asm volatile("" : "+v"(off));
    return off;
}

// Element of slot k in a direction array: k, or in the half-interleaved
// 16-bit layout (hs = N / 2 > 0) 2k for k < hs and 2(k - hs) + 1 above, so
// that slots k and k + hs (nodes half the last axis apart) share a word.
__device__ __forceinline__ unsigned slot_elem(unsigned k, unsigned hs)
{
    return hs == 0u ? k : (k < hs ? 2u * k : 2u * (k - hs) + 1u);
}

// index-shift gather of node n's populations (solver.cu:77-94) from the
// slots n + dd[0][e_x+1] + dd[1][e_y+1] + dd[2][e_z+1]
template <class L, class St, bool WIDE>
__device__ __forceinline__ void gather_node(const StepArgs<typename St::C> &a, int n, const int (&dd)[3][3],
                                            typename St::C (&v)[L::Q])
{
    using S = typename St::S;
    if constexpr (WIDE) {
        if (a.wide_rows) {
            wide_row_slots<L, St>(a, n, dd, [&](auto ii, const S *p) { v[decltype(ii)::value] = St::load(ldg(p)); });
            return;
        }
    }
    const S *const fb = (const S *)a.f;
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const int k = n + dd[0][L::ex[i] + 1] + dd[1][L::ey[i] + 1] + dd[2][L::ez[i] + 1];
        if constexpr (WIDE) {
            v[i] = St::load(ldg(fb + (size_t)i * a.npad + k));
        } else {
            const char *base = (const char *)(fb + (size_t)i * a.npad);
            const unsigned e = sizeof(S) == 2 ? slot_elem((unsigned)k, a.hs) : (unsigned)k;
            v[i] = St::load(ldg((const S *)(base + e * (unsigned)sizeof(S))));
        }
    });
}

// write-back to the same slots (solver.cu:119-123); the 16-bit formats
// round stochastically (storage.hpp)
template <class L, class St, bool WIDE>
__device__ __forceinline__ void scatter_node(const StepArgs<typename St::C> &a, int n, const int (&dd)[3][3],
                                             const typename St::C (&v)[L::Q])
{
    using S = typename St::S;
    S *const fb = (S *)a.f;
    unsigned r = 0;
    if constexpr (St::stochastic) {
        // the call's seed from its shift counters s_d = t mod L_d: one value
        // per t mod lcm(L), so a graph replay, a direct launch and a batch
        // member draw the same words for the same call
        const unsigned ts = (unsigned)a.s[0] + (unsigned)a.L[0] * ((unsigned)a.s[1] + (unsigned)a.L[1] * (unsigned)a.s[2]);
        r = sr_seed(n, ts);
    }
    if constexpr (WIDE) {
        if (a.wide_rows) {
            wide_row_slots<L, St>(a, n, dd, [&](auto ii, S *p) {
                constexpr int i = decltype(ii)::value;
                if constexpr (St::stochastic) {
                    stg(p, St::store_sr(v[i], r));
                    r = sr_next(r);
                } else {
                    stg(p, St::store(v[i]));
                }
            });
            return;
        }
    }
    sfor<0, L::Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const int k = n + dd[0][L::ex[i] + 1] + dd[1][L::ey[i] + 1] + dd[2][L::ez[i] + 1];
        S s;
        if constexpr (St::stochastic) {
            s = St::store_sr(v[i], r);
            r = sr_next(r);
        } else {
            s = St::store(v[i]);
        }
        if constexpr (WIDE) {
            stg(fb + (size_t)i * a.npad + k, s);
        } else {
            char *base = (char *)(fb + (size_t)i * a.npad);
            const unsigned e = sizeof(S) == 2 ? slot_elem((unsigned)k, a.hs) : (unsigned)k;
            stg((S *)(base + lane_offset(e * (unsigned)sizeof(S))), s);
        }
    });
}

// gather, update, scatter
template <class L, class St, int FAM, bool GUO, bool SMAG, bool MFIX, bool WIDE>
__device__ __forceinline__ void step_node(const StepArgs<typename St::C> &a, int n, const int (&dd)[3][3])
{
    typename St::C v[L::Q];
    constexpr bool PRIO = St::gather_prio && L::D == 3;
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
    gather_node<L, St, WIDE>(a, n, dd, v);
    if constexpr (PRIO) {
        // keep the loads above the priority drop
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
    }
    node_update<L, St, FAM, GUO, SMAG, MFIX>(a, n, v);
    scatter_node<L, St, WIDE>(a, n, dd, v);
}

// Slot offsets of node n for this call: the slot of direction i is
// n + dd[0][e_x+1] + dd[1][e_y+1] + dd[2][e_z+1], the reference's
// ((c - t e) mod L) per axis (solver.cu:77-92) relative to the node's own
// slot, wrap included.
template <class L, class C>
__device__ __forceinline__ void slot_offsets(const StepArgs<C> &a, int n, int (&dd)[3][3])
{
    const int Lx = a.L[0], Ly = a.L[1];
    const int row = fast_div(n, a.mx, a.lx);
    const int x = n - row * Lx;
    const int z = fast_div(row, a.my, a.ly);
    const int y = row - z * Ly;
    const int c3[3] = {x, y, z};
    sfor<0, L::D>([&](auto aa) {
        constexpr int ax = decltype(aa)::value;
        const int c = c3[ax];
        dd[ax][1] = 0;
        dd[ax][2] = (c >= a.s[ax]) ? a.dp[ax] : a.dpw[ax];
        dd[ax][0] = (c + a.s[ax] < a.L[ax]) ? a.dm[ax] : a.dmw[ax];
    });
    if constexpr (L::D == 2) {
        dd[2][0] = dd[2][1] = dd[2][2] = 0;
    }
}

// One thread per node, nodes in slot order (x fastest).  For a fixed
// direction, consecutive nodes read consecutive slots of the source row
// (rotated by t*e_x), so every wave's accesses are contiguous apart from
// at most one wrap point.
template <class L, class St, int FAM, bool GUO, bool SMAG, bool MFIX, bool WIDE>
__global__ __launch_bounds__(kBlock) void step_kernel(const StepArgs<typename St::C> a)
{
    // grid <= 2^23 + 2^5 blocks (row_grid): the node index is exact in unsigned
    const unsigned nu = row_chunk(blockIdx.x, a.rlog) * (unsigned)kBlock + threadIdx.x;
    if (nu >= (unsigned)a.nnodes) return;
    const int n = (int)nu;
    int dd[3][3];
    slot_offsets<L>(a, n, dd);
    // WIDE = false (direction arrays below 4 GiB): a uniform base per
    // direction plus a 32-bit byte offset per lane, no 64-bit address
    // arithmetic per access
    step_node<L, St, FAM, GUO, SMAG, MFIX, WIDE>(a, n, dd);
}

// Two nodes per lane for the 16-bit formats on calls whose x shift
// s_x = t mod Lx is even, with Lx even.  Nodes n = 2p and n + 1 lie in one
// row and have the same slot offsets in every direction (a wrap between x
// and x + 1 would need x + s_x odd), and slot n + offset is even, so each
// direction's two populations are one aligned 32-bit word: one load and one
// store per direction for both nodes, half the memory instructions of
// step_kernel.  Each node takes step_kernel's update and stochastic-rounding
// words, so the results are bit-identical.  Direction arrays below 4 GiB
// only.  (FP32 pairs in 64-bit words are correct too but 2-3 % slower than
// one node per lane at every residency cap, so FP32 keeps step_kernel.)
template <class L, class St, int FAM, bool GUO, bool SMAG, bool MFIX>
__global__ __launch_bounds__(kBlock) void step_pair_kernel(const StepArgs<float> a)
{
    using S = typename St::S;
    static_assert(sizeof(S) == 2, "pairs of 16-bit slots");
    constexpr int Q = L::Q;
    const int n = 2 * ((int)blockIdx.x * kBlock + (int)threadIdx.x);
    if (n >= a.nnodes) return;
    int dd[3][3];
    slot_offsets<L>(a, n, dd);
    const S *const fb = (const S *)a.f;
    float v0[Q], v1[Q];
    constexpr bool PRIO = St::gather_prio && L::D == 3;
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
    sfor<0, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const int k = n + dd[0][L::ex[i] + 1] + dd[1][L::ey[i] + 1] + dd[2][L::ez[i] + 1];
        const char *base = (const char *)(fb + (size_t)i * a.npad);
        const unsigned w = ldg((const unsigned *)(base + (unsigned)k * 2u));
        typedef S s2 __attribute__((ext_vector_type(2)));
        const s2 wv = __builtin_bit_cast(s2, w);
        v0[i] = St::load(wv.x);
        v1[i] = St::load(wv.y);
    });
    if constexpr (PRIO) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
    }
    node_update<L, St, FAM, GUO, SMAG, MFIX>(a, n, v0);
    node_update<L, St, FAM, GUO, SMAG, MFIX>(a, n + 1, v1);
    const unsigned ts = (unsigned)a.s[0] + (unsigned)a.L[0] * ((unsigned)a.s[1] + (unsigned)a.L[1] * (unsigned)a.s[2]);
    unsigned r0 = sr_seed(n, ts), r1 = sr_seed(n + 1, ts);
    S *const fw = (S *)a.f;
    sfor<0, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        const int k = n + dd[0][L::ex[i] + 1] + dd[1][L::ey[i] + 1] + dd[2][L::ez[i] + 1];
        const unsigned w = St::store_sr_pair(v0[i], r0, v1[i], r1);
        r0 = sr_next(r0);
        r1 = sr_next(r1);
        char *base = (char *)(fw + (size_t)i * a.npad);
        stg((unsigned *)(base + (unsigned)k * 2u), w);
    });
}

// Two nodes per lane in the half-interleaved 16-bit layout (StepArgs::hs =
// N / 2 > 0), on every call: lane p takes nodes n0 = p and n1 = p + hs,
// which sit half the last axis apart (z and z + Lz/2; y and y + Ly/2 on
// D2Q9) at the same x and y.  The index shift moves both by the same
// vector, so for every direction their slots are k0 and k0 +- hs: element
// 2(k0 mod hs) and the element next to it, one aligned 32-bit word, with
// node n0 in the low half when k0 < hs and in the high half otherwise.  One
// dword load and one dword store per direction for both nodes, on odd and
// even shifts alike; consecutive lanes are consecutive x, so a wave's
// accesses for a direction are 64 consecutive words (one rotation point),
// the FP32 step's pattern.  Each node takes step_kernel's update and
// stochastic-rounding words: bit-identical results.  Direction arrays below
// 4 GiB only.
//
// Which half is n0's depends on the last axis alone (the other axes' slot
// offsets stay within n0's plane or row, which lies in the lower half):
// always the low half for directions with no last-axis component, else one
// flag per sign of that component.  The flags fold into the last-axis word
// offsets and into a byte selector (v_perm_b32) that swaps the halves where
// n0 holds the high one, on the loaded and on the stored word alike.
template <class L, class St, int FAM, bool GUO, bool SMAG, bool MFIX>
__global__ __launch_bounds__(kBlock) void step_half_kernel(const StepArgs<float> a)
{
    using S = typename St::S;
    static_assert(sizeof(S) == 2, "pairs of 16-bit slots");
    constexpr int Q = L::Q;
    constexpr int A = L::D - 1;  // the split axis
    // grid = row_grid(ceil(hs / 256)) blocks: p < 2^31
    const unsigned p = row_chunk(blockIdx.x, a.rlog) * (unsigned)kBlock + threadIdx.x;
    const unsigned hs = a.hs;
    if (p >= hs) return;
    const int n0 = (int)p, n1 = (int)(p + hs);
    int dd[3][3];
    slot_offsets<L>(a, n0, dd);
    // word offsets along the split axis, [e + 1] for component e, and the
    // half selectors of the two signs
    constexpr unsigned keep = 0x03020100u, swap = 0x01000302u;
    const bool lowp = (unsigned)(n0 + dd[A][2]) < hs, lowm = (unsigned)(n0 + dd[A][0]) < hs;
    int dw[3];
    dw[1] = 0;
    dw[2] = lowp ? dd[A][2] : dd[A][2] - (int)hs;
    dw[0] = lowm ? dd[A][0] : dd[A][0] - (int)hs;
    const unsigned selp = lowp ? keep : swap, selm = lowm ? keep : swap;
    float v0[Q], v1[Q];
    constexpr bool PRIO = St::gather_prio && L::D == 3;
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
    sfor<0, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int e = A == 2 ? L::ez[i] : L::ey[i];
        const int other = A == 2 ? dd[0][L::ex[i] + 1] + dd[1][L::ey[i] + 1] : dd[0][L::ex[i] + 1];
        const unsigned w = (unsigned)(n0 + other + dw[e + 1]);
        const char *base = (const char *)((const S *)a.f + (size_t)i * a.npad);
        unsigned word = ldg((const unsigned *)(base + w * 4u));
        if constexpr (e != 0) word = __builtin_amdgcn_perm(word, word, e > 0 ? selp : selm);
        typedef S s2 __attribute__((ext_vector_type(2)));
        const s2 wv = __builtin_bit_cast(s2, word);
        v0[i] = St::load(wv.x);
        v1[i] = St::load(wv.y);
    });
    if constexpr (PRIO) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
    }
    node_update<L, St, FAM, GUO, SMAG, MFIX>(a, n0, v0);
    node_update<L, St, FAM, GUO, SMAG, MFIX>(a, n1, v1);
    const unsigned ts = (unsigned)a.s[0] + (unsigned)a.L[0] * ((unsigned)a.s[1] + (unsigned)a.L[1] * (unsigned)a.s[2]);
    unsigned r0 = sr_seed(n0, ts), r1 = sr_seed(n1, ts);
    sfor<0, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        constexpr int e = A == 2 ? L::ez[i] : L::ey[i];
        const int other = A == 2 ? dd[0][L::ex[i] + 1] + dd[1][L::ey[i] + 1] : dd[0][L::ex[i] + 1];
        const unsigned w = (unsigned)(n0 + other + dw[e + 1]);
        // node n0's value in the low half, then the halves swapped where n0
        // holds the high one
        unsigned word = St::store_sr_pair(v0[i], r0, v1[i], r1);
        r0 = sr_next(r0);
        r1 = sr_next(r1);
        if constexpr (e != 0) word = __builtin_amdgcn_perm(word, word, e > 0 ? selp : selm);
        char *base = (char *)((S *)a.f + (size_t)i * a.npad);
        stg((unsigned *)(base + lane_offset(w * 4u)), word);
    });
}

// Gathers the post-stream populations of the neighbours of the sf nodes
// before the step kernel overwrites any slot.  sfnodes[k] is a node index,
// sfnb[k] its neighbour's index.
template <class L, class St>
__global__ __launch_bounds__(kBlock) void sf_gather_kernel(const StepArgs<typename St::C> a,
                                                           const int *sfnb, int nsf, typename St::C *out)
{
    using C = typename St::C;
    using S = typename St::S;
    constexpr int Q = L::Q;
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= nsf) return;
    const int Lx = a.L[0], Ly = a.L[1], Lz = a.L[2];
    const int nb = sfnb[k];
    const int x = nb % Lx;
    const int y = (nb / Lx) % Ly;
    const int z = nb / (Lx * Ly);
    const S *fb = (const S *)a.f;
    sfor<0, Q>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        int xs = x - L::ex[i] * a.s[0];
        xs = xs < 0 ? xs + Lx : (xs >= Lx ? xs - Lx : xs);
        int ys = y - L::ey[i] * a.s[1];
        ys = ys < 0 ? ys + Ly : (ys >= Ly ? ys - Ly : ys);
        int zs = z - L::ez[i] * a.s[2];
        zs = zs < 0 ? zs + Lz : (zs >= Lz ? zs - Lz : zs);
        const int id = (zs * Ly + ys) * Lx + xs;
        const unsigned e = sizeof(S) == 2 ? slot_elem((unsigned)id, a.hs) : (unsigned)id;
        C v = St::load(fb[(size_t)i * a.npad + e]);
        if constexpr (St::shifted) v = v + C(L::w[i]);
        out[(size_t)k * Q + i] = v;
    });
}

}  // namespace lbmc
