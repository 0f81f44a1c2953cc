// Node index arithmetic of the kernels that walk a lattice in tiles of
// `block` consecutive nodes (the closure step's grid-stride loop, the batch
// kernels' chunks).  A lattice holds at most 2^31 - 1 nodes (lbmc_create),
// so a tile start plus a lane may pass INT_MAX by up to block - 1 in the
// last, partial tile, and a chunk's tiles by up to the chunk length: the
// sums are formed in unsigned arithmetic, which holds them exactly (they
// stay below 2^32), and compared with the node count before a live index is
// narrowed to int.  Plain constexpr functions: the same code runs in the
// kernels and in the host-side limit test.
#pragma once

namespace lbmc {

// tiles of `block` nodes covering n nodes, 0 <= n < 2^31
constexpr unsigned tile_count(int n, unsigned block) { return ((unsigned)n + block - 1u) / block; }

// node of lane `lane` in tile `tile` of a chunk whose first node is `first`
constexpr unsigned tile_node(unsigned first, unsigned tile, unsigned block, unsigned lane)
{
    return first + tile * block + lane;
}

// Block order of the one-block-per-chunk step kernels on lattices whose
// x-rows hold R >= 2 chunks of 256 nodes.  The index shift rotates a row by
// t in every direction with e_x != 0, so unless t is a multiple of 32 the
// cache line at each chunk boundary inside a row holds slots of both
// chunks; blocks are dealt to the 8 XCDs round robin (block b to b mod 8),
// so in the plain order the two chunks sit on two XCDs and both L2s fetch
// and write back that line (512^3 D3Q15 FP32: 65.0 read + 61.3 written
// bytes per node against 60 + 60).  row_chunk gives block b chunk
// c = ((g * 8 + b mod 8) << rlog) + j, g = b >> (rlog + 3), j = (b >> 3)
// mod 2^rlog: 2^rlog consecutive chunks on one XCD, the next 2^rlog on the
// next XCD.  The grid covers whole groups of 8 * 2^rlog blocks; blocks past
// the last chunk get chunks past the last node.
constexpr unsigned row_chunk(unsigned b, unsigned rlog)
{
    return rlog == 0u ? b : ((((b >> (rlog + 3u)) << 3u) + (b & 7u)) << rlog) + ((b >> 3u) & ((1u << rlog) - 1u));
}

// blocks of a launch over `chunks` chunks of 256 nodes (a multiple of
// 2^rlog) in row_chunk order
constexpr unsigned row_grid(unsigned chunks, unsigned rlog)
{
    return rlog == 0u ? chunks : (((chunks >> rlog) + 7u) / 8u * 8u) << rlog;
}

}  // namespace lbmc
