// Batched step kernels with StFP64 population storage (built with
// contraction off, like step_fp64.hip, so members round like the reference).
#include "batch_kernel.hpp"
#include "dispatch.hpp"

namespace lbmc {

hipError_t launch_batch_fp64(int lattice, const void *members, const int2 *chunks, int nchunks, int grid, int sub,
                             int lds, hipStream_t s)
{
    return launch_batch_st<StFP64>(lattice, members, chunks, nchunks, BatchShape{grid, sub, lds},
                                   BatchEval{nullptr, 0, 0, nullptr, nullptr, false}, s);
}

}  // namespace lbmc
