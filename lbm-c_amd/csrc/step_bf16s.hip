// Fused step kernels with StBF16S population storage.
#include "step_launch.hpp"

namespace lbmc {

hipError_t launch_step_bf16s(const KernelKey &k, const StepArgs<float> &a, int grid, hipStream_t s)
{
    return launch_step_st<StBF16S>(k, a, grid, s);
}

hipError_t launch_sf_gather_bf16s(int lattice, const StepArgs<float> &a, const int *sfnb, int nsf,
                                  float *out, hipStream_t s)
{
    return launch_sf_st<StBF16S>(lattice, a, sfnb, nsf, out, s);
}

}  // namespace lbmc
