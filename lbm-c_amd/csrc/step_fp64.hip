// Fused step kernels with StFP64 population storage.
#include "step_launch.hpp"

namespace lbmc {

hipError_t launch_step_fp64(const KernelKey &k, const StepArgs<double> &a, int grid, hipStream_t s)
{
    return launch_step_st<StFP64>(k, a, grid, s);
}

hipError_t launch_sf_gather_fp64(int lattice, const StepArgs<double> &a, const int *sfnb, int nsf,
                                  double *out, hipStream_t s)
{
    return launch_sf_st<StFP64>(lattice, a, sfnb, nsf, out, s);
}

}  // namespace lbmc
