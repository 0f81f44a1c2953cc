// Fused step kernels with StFP32 population storage.
#include "step_launch.hpp"

namespace lbmc {

hipError_t launch_step_fp32(const KernelKey &k, const StepArgs<float> &a, int grid, hipStream_t s)
{
    return launch_step_st<StFP32>(k, a, grid, s);
}

hipError_t launch_sf_gather_fp32(int lattice, const StepArgs<float> &a, const int *sfnb, int nsf,
                                  float *out, hipStream_t s)
{
    return launch_sf_st<StFP32>(lattice, a, sfnb, nsf, out, s);
}

}  // namespace lbmc
