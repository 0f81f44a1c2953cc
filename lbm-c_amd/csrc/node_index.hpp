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

}  // namespace lbmc
