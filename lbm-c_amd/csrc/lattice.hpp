// Compile-time lattice definitions for the fused step kernels.
//
// Velocity sets, weights, opposite directions and the MRT transform follow
// LBM-C/macros.cu:14-58 (D2Q9) and :60-125 (D3Q15).  Everything here is
// constexpr so that the kernels fold every multiply by 0 or +-1 away; the
// remaining operations keep the reference's order.
#pragma once

#include <cstdint>

namespace lbmc {

struct D2Q9 {
    static constexpr int D = 2;
    static constexpr int Q = 9;
    static constexpr int id = 0;
    static constexpr int ex[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int ey[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int ez[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    static constexpr double w[Q] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int M[Q][Q] = {
        {1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2},
        {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
        {0, -2, 0, 2, 0, 1, -1, -1, 1},  {0, 0, 1, 0, -1, 1, 1, -1, -1},
        {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 1, -1, 1, -1}};
    // M^-1 exactly as the reference tabulates it (macros.cu:31-43): the
    // values (including the 1.39e-17 entries) are data that bitwise parity
    // depends on.
    static constexpr double Minv[Q][Q] = {
        {0.1111111111111111, -0.11111111111111112, 0.1111111111111111, 0, 0, 0, 0, 0, 0},
        {0.1111111111111111, -0.02777777777777779, -0.055555555555555566, 0.16666666666666666,
         -0.16666666666666669, 0, 0, 0.25, 0},
        {0.1111111111111111, -0.027777777777777762, -0.055555555555555539, 0,
         -1.3877787807814457e-017, 0.16666666666666666, -0.16666666666666669, -0.25, 0},
        {0.1111111111111111, -0.02777777777777779, -0.055555555555555566, -0.16666666666666666,
         0.16666666666666669, 0, 0, 0.25, 0},
        {0.1111111111111111, -0.02777777777777779, -0.055555555555555566, 0,
         1.3877787807814457e-017, -0.16666666666666666, 0.16666666666666669, -0.25, 0},
        {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, 0.16666666666666666,
         0.083333333333333329, 0.16666666666666666, 0.083333333333333329, 0, 0.25},
        {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, -0.16666666666666666,
         -0.083333333333333329, 0.16666666666666666, 0.083333333333333329, 0, -0.25},
        {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, -0.16666666666666666,
         -0.083333333333333329, -0.16666666666666666, -0.083333333333333329, 0, 0.25},
        {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, 0.16666666666666666,
         0.083333333333333329, -0.16666666666666666, -0.083333333333333329, 0, -0.25}};
};

struct D3Q15 {
    static constexpr int D = 3;
    static constexpr int Q = 15;
    static constexpr int id = 1;
    static constexpr int ex[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
    static constexpr int ey[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int ez[Q] = {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};
    static constexpr double w[Q] = {2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                    1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                                    1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
    static constexpr int M[Q][Q] = {
        {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
        {-2, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1},
        {16, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1},
        {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1},
        {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1},
        {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
        {0, 0, 0, -4, 4, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
        {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1},
        {0, 0, 0, 0, 0, -4, 4, 1, 1, 1, 1, -1, -1, -1, -1},
        {0, 2, 2, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
        {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, -1, -1, 1, 1},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1},
        {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1}};
    // squared row norms of M: M has orthogonal rows, so M^-1 = M^T diag(1/nrm)
    static constexpr int nrm[Q] = {15, 18, 360, 10, 40, 10, 40, 10, 40, 12, 4, 8, 8, 8, 8};
};

// e component by axis
template <class L>
__host__ __device__ constexpr int e_of(int d, int i)
{
    return d == 0 ? L::ex[i] : (d == 1 ? L::ey[i] : L::ez[i]);
}

// M^-1 entry.  D2Q9: the reference's table.  D3Q15: the reference's table is
// all zeros (integer fractions, macros.cu:99-113); with FIX the table those
// fractions spell, M^T diag(1/nrm).
template <class L, bool FIX>
struct MinvOf;
template <bool FIX>
struct MinvOf<D2Q9, FIX> {
    static constexpr double at(int i, int j) { return D2Q9::Minv[i][j]; }
};
template <bool FIX>
struct MinvOf<D3Q15, FIX> {
    static constexpr double at(int i, int j)
    {
        return FIX ? (double)D3Q15::M[j][i] / (double)D3Q15::nrm[j] : 0.0;
    }
};

// macro BC code 1..15 -> prescribed set, listing order of macro_conditions[]
// (LBM-C/boundary_conditions/macro_bc.cu:12-17); bit 0 rho, bits 1..3 u_x..u_z
// (codes 8..15 add rho to the velocity sets of codes 0..7; the sets of
// codes 0..7 are packed one per nibble so no table lives in memory)
__host__ __device__ constexpr unsigned macro_set(int code)
{
    return ((code & 8) ? 1u : 0u) | ((0xECA68420u >> (4 * (code & 7))) & 0xFu);
}

// node class descriptor bits (the 1-byte mask selects one of these per node)
enum : uint16_t {
    CLS_SOLID = 0x1,
    CLS_POROUS = 0x2,
};
__host__ __device__ inline int cls_macro(uint16_t d) { return (d >> 4) & 15; }
__host__ __device__ inline int cls_micro(uint16_t d) { return (d >> 8) & 15; }

}  // namespace lbmc
