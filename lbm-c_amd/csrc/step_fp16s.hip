// Fused step kernels with StFP16S population storage.
#include "step_launch.hpp"

namespace lbmc {

hipError_t launch_step_fp16s(const KernelKey &k, const StepArgs<float> &a, int grid, hipStream_t s)
{
    return launch_step_st<StFP16S>(k, a, grid, s);
}

hipError_t launch_sf_gather_fp16s(int lattice, const StepArgs<float> &a, const int *sfnb, int nsf,
                                  float *out, hipStream_t s)
{
    return launch_sf_st<StFP16S>(lattice, a, sfnb, nsf, out, s);
}

}  // namespace lbmc
