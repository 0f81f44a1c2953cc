// Node index arithmetic of the step, closure and batch kernels at the node
// limit (2^31 - 1 nodes per lattice, lbmc_create), run on the host under
// UndefinedBehaviorSanitizer: the kernels' loop structure is replayed for
// the tiles near the end of the index range, every live index must lie in
// [first node of its chunk, N), and the last node must be covered.
//
// The closure and batch tiles use lbmc::tile_count / lbmc::tile_node, the
// one-node step kernels lbmc::row_chunk / lbmc::row_grid (the header the
// kernels include).  The two-node step kernel indexes with plain int
// arithmetic (its grid ends exactly at the last node); it is restated here
// so that a signed overflow in it would trap.
//
// Build: g++ -std=c++17 -fsanitize=undefined,address -fno-sanitize-recover=all
// With -DSIGNED_TILES the closure loop is replayed with the int arithmetic it
// used before (n0 = b * 256 + lane), which overflows near the limit: the
// sanitizer must stop that build (the test's check that this one bites).
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <vector>

#include "../../lbm-c_amd/csrc/node_index.hpp"

namespace {

constexpr int kBlock = 256;
int failures = 0;

void check(bool ok, const char *what, long long n, long long N)
{
    if (!ok) {
        std::printf("FAIL %s: index %lld, N %lld\n", what, n, N);
        failures++;
    }
}

// closure_step_kernel: blocks stride over the positions of the 256-node
// chunks in row_chunk order (rlog 0: the chunks themselves), lanes past the
// last node evaluate a copy of node N - 1 and write nothing
void closure_tiles(int N, unsigned grid, unsigned rlog = 0)
{
#ifdef SIGNED_TILES
    const int nchunks = (N + kBlock - 1) / kBlock;
    const int npos = nchunks;
#else
    const unsigned nchunks = lbmc::tile_count(N, kBlock);
    const unsigned npos = lbmc::row_grid(nchunks, rlog);
#endif
    check((long long)nchunks * kBlock >= N && ((long long)nchunks - 1) * kBlock < N, "closure chunk count",
          nchunks, N);
    bool last = false;
    for (unsigned blk = 0; blk < grid; blk++) {
        // the block's last few positions only: the ones that approach N
        unsigned p = blk;
        if ((unsigned)npos > 4 * grid) p += ((unsigned)npos / grid - 3) * grid;
        for (; p < (unsigned)npos; p += grid) {
#ifdef SIGNED_TILES
            const unsigned b = p;
#else
            const unsigned b = lbmc::row_chunk(p, rlog);
            if (b >= nchunks) continue;
#endif
            for (unsigned lane = 0; lane < (unsigned)kBlock; lane++) {
#ifdef SIGNED_TILES
                const int n0 = (int)b * kBlock + (int)lane;
                const bool live = n0 < N;
                const int n = live ? n0 : N - 1;
#else
                const unsigned n0 = lbmc::tile_node(0u, b, kBlock, lane);
                const bool live = n0 < (unsigned)N;
                const int n = live ? (int)n0 : N - 1;
#endif
                check(n >= 0 && n < N, "closure node", n, N);
                if (live && n == N - 1) last = true;
            }
        }
    }
    check(last, "closure covers the last node", N - 1, N);
}

// batch_step_kernel / batch_eval_kernel: chunks of chunk_nodes nodes of one
// member, `sub` tiles each
void batch_tiles(int N, int chunk_nodes)
{
    const int sub = chunk_nodes / kBlock;
    bool last = false;
    long long first_chunk = ((long long)N - 1) / chunk_nodes * chunk_nodes;
    for (long long first = first_chunk - 2LL * chunk_nodes; first < N; first += chunk_nodes) {
        if (first < 0) continue;
        for (int j = 0; j < sub; j++) {
            const unsigned t0 = lbmc::tile_node((unsigned)first, (unsigned)j, kBlock, 0u);
            if (t0 >= (unsigned)N) break;
            for (unsigned lane = 0; lane < (unsigned)kBlock; lane++) {
                const unsigned n0 = t0 + lane;
                if (n0 >= (unsigned)N) continue;
                const int n = (int)n0;
                check(n >= first && n < N, "batch node", n, N);
                if (n == N - 1) last = true;
            }
        }
    }
    check(last, "batch covers the last node", N - 1, N);
}

// step_kernel, step_half_kernel, closure_reuse_kernel: grid =
// row_grid(ceil(N / 256), rlog), n = row_chunk(b, rlog) * 256 + lane in
// unsigned arithmetic (rlog > 0 only where N is a multiple of 256 << rlog)
void step_tiles(int N, unsigned rlog)
{
    const unsigned chunks = lbmc::tile_count(N, kBlock);
    const unsigned grid = lbmc::row_grid(chunks, rlog);
    check(grid >= chunks && grid - chunks < (8u << rlog), "step grid", grid, N);
    bool last = false;
    const unsigned tail = 16u << rlog;
    for (unsigned b = grid > tail ? grid - tail : 0; b < grid; b++) {
        const unsigned c = lbmc::row_chunk(b, rlog);
        for (unsigned lane = 0; lane < (unsigned)kBlock; lane++) {
            const unsigned nu = c * (unsigned)kBlock + lane;
            if (nu >= (unsigned)N) continue;
            const int n = (int)nu;
            check(n >= 0, "step node", n, N);
            if (n == N - 1) last = true;
        }
    }
    check(last, "step covers the last node", N - 1, N);
}

// the row order over a whole small lattice: every node exactly once, and
// the 2^rlog chunks of a group on one XCD (blocks are dealt b mod 8)
void row_order_cover(int N, unsigned rlog)
{
    const unsigned chunks = lbmc::tile_count(N, kBlock);
    const unsigned grid = lbmc::row_grid(chunks, rlog);
    std::vector<unsigned char> seen((size_t)N, 0);
    for (unsigned b = 0; b < grid; b++) {
        const unsigned c = lbmc::row_chunk(b, rlog);
        if (c < chunks) check((c >> rlog) % 8u == b % 8u, "row group on one XCD", c, N);
        for (unsigned lane = 0; lane < (unsigned)kBlock; lane++) {
            const unsigned nu = c * (unsigned)kBlock + lane;
            if (nu >= (unsigned)N) continue;
            check(seen[nu] == 0, "node taken twice", nu, N);
            seen[nu] = 1;
        }
    }
    for (int n = 0; n < N; n++)
        if (!seen[(size_t)n]) {
            check(false, "node not taken", n, N);
            break;
        }
}

// step_pair_kernel (N even): grid = ceil(N / 2 / 256), n = 2 (b * 256 + lane)
void pair_tiles(int N)
{
    const int grid = (N / 2 + kBlock - 1) / kBlock;
    bool last = false;
    for (int b = grid - 3 < 0 ? 0 : grid - 3; b < grid; b++)
        for (int lane = 0; lane < kBlock; lane++) {
            const int n = 2 * (b * kBlock + lane);
            if (n >= N) continue;
            check(n >= 0 && n + 1 < N, "pair node", n, N);
            if (n + 1 == N - 1) last = true;
        }
    check(last, "pairs cover the last node", N - 1, N);
}

}  // namespace

int main()
{
    const int limit = 0x7fffffff;  // 2^31 - 1
    const int sizes[] = {limit, limit - 1, limit - 255, limit - 256, limit - 257, limit - 16383, 1 << 30, 300, 1};
    for (int N : sizes) {
        for (unsigned grid : {1u, 7u, 2048u}) closure_tiles(N, grid);
        for (int cn : {256, 512, 768, 1024, 2048, 16384}) batch_tiles(N, cn);
        step_tiles(N, 0);
        if (N % 2 == 0) pair_tiles(N);
    }
    // row order: lattices of 512- and 1024-node rows near the limit and small
    for (unsigned rlog : {1u, 2u}) {
        const int row = kBlock << rlog;
        for (int N : {limit / row * row, (limit / row - 7) * row, (1 << 30), 30 * row, 15 * row, row}) {
            step_tiles(N, rlog);
            for (unsigned grid : {8u, 2048u, 8192u}) closure_tiles(N, grid, rlog);
            if (N <= (1 << 20)) row_order_cover(N, rlog);
        }
    }
    std::printf("%s\n", failures ? "index limits: FAILED" : "index limits: ok");
    return failures ? 1 : 0;
}
