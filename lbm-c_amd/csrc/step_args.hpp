// Kernel argument block shared by the host dispatcher and the step kernels.
#pragma once

#include <cstdint>

namespace lbmc {

// threads per block of the step kernels
constexpr int kBlock = 256;

// Division by a fixed divisor d >= 1 for numerators n < 2^31:
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, n / d = (umulhi(n, m) + n) >> l.
struct FastDiv {
    unsigned m;
    int l;
};
inline FastDiv make_fastdiv(int d)
{
    int l = 0;
    while ((1LL << l) < d) l++;
    const unsigned long long m = ((1ULL << 32) * ((1ULL << l) - (unsigned long long)d)) / (unsigned long long)d + 1;
    return FastDiv{(unsigned)m, l};
}

// which fused kernel variant a launch needs
struct KernelKey {
    int lattice;  // 0 D2Q9, 1 D3Q15
    int family;   // 0 BGK (BGK and NTPOR), 1 MRT (MRT and MRTPOR)
    int guo;      // Guo body force
    int smag;     // Smagorinsky (BGK family only)
    int mfix;     // corrected D3Q15 M^-1 (MRT family only)
    int wide;     // a direction array spans 4 GiB or more: 64-bit offsets
    int closure;  // learned relaxation-time closure (BGK family, FP32 compute)
    int closure_fp16;  // its network with fp16 operands (else bf16)
    int lds;      // dynamic LDS bytes per block of the step kernels (which use
                  //   none): caps the blocks resident per CU, 0 = no cap
    int pairs;    // two nodes per lane: on every call in the half-interleaved 16-bit layout
                  //   (StepArgs::hs), else on calls with an even x shift
};

enum : unsigned {
    QK_FIX_D3Q15_MINV = 0x1u,
    QK_FIX_D3Q15_MEQ = 0x2u,
    QK_FIX_D3Q15_ZH = 0x4u,
};

template <class C>
struct StepArgs {
    void *f;                 // Q directions x npad slots of the storage type
    long long npad;          // elements between direction arrays
    const uint8_t *mask;     // node class per node
    const uint16_t *cls;     // class descriptors (256 entries)
    const C *porosity;       // per node; read at porous classes only
    const C *force[3];       // per-node force, or null for a uniform force
    C F[3];                  // uniform force
    C *rho;                  // macro arrays: output, and prescribed values
    C *u[3];                 //   at macro-BC nodes (LBM-C/boundary_conditions/macro_bc.cu)
    const C *sfbuf;          // gathered neighbour populations for sf nodes
    const int *sfidx;        // per node: row of sfbuf (sf nodes only)
    int L[3];
    int s[3];                // t mod L[d]: the index-shift of this call
    int nnodes;              // Lx * Ly * Lz
    unsigned mx, my;         // magic multipliers: n / Lx and row / Ly
    int lx, ly;              //   (q = (umulhi(n, m) + n) >> l, n < 2^31)
    int wide_rows;           // 64-bit-offset kernels with Lx % 64 == 0: every wave lies in
                             //   one x-row, so each direction's slot row is wave-uniform
    unsigned rlog;           // blocks take 2^rlog consecutive 256-node chunks per XCD (row_chunk), 0: plain order
    unsigned hs;             // 16-bit storage in the half-interleaved layout: hs = nnodes / 2
                             //   (the halves of the last axis); slot k < hs is element 2k and
                             //   slot k + hs element 2k + 1 of the direction array, so the two
                             //   share one 32-bit word.  0: slot k is element k
    // slot offsets of this call relative to the node index: for e = +1 along
    // an axis the slot is the node shifted by -s (or by L - s once wrapped),
    // for e = -1 by +s (or s - L); scaled by the axis stride
    int dp[3], dpw[3];       // e = +1: -s*stride, (L - s)*stride
    int dm[3], dmw[3];       // e = -1: +s*stride, (s - L)*stride
    C tau;
    C c_smag;
    C tau_mrt[15];
    int store;               // store rho/u this call
    int porous;              // NTPOR/MRTPOR with porous nodes present
    int all_fluid;           // every node is plain fluid: class 0 without reading the mask
    unsigned quirks;
    const void *mlp;         // closure weights (MlpImage) when the closure is on
    float *tcache;           // closure: per-node tau of the last evaluation (NaN: fall back)
    int tmode;               // closure: 0 evaluate, 1 evaluate and record, 2 use the record
    C closure_shift;         // closure: added to the network's tau'
};

}  // namespace lbmc
