// Host-side launchers of the device kernels (one translation unit per
// population storage format, so that the FP64 unit can be built with
// contraction off).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lbmc.h"
#include "step_args.hpp"

namespace lbmc {

// fused step, step_<storage>.hip
hipError_t launch_step_fp64(const KernelKey &k, const StepArgs<double> &a, int grid, hipStream_t s);
hipError_t launch_step_fp32(const KernelKey &k, const StepArgs<float> &a, int grid, hipStream_t s);
hipError_t launch_step_fp16s(const KernelKey &k, const StepArgs<float> &a, int grid, hipStream_t s);
hipError_t launch_step_bf16s(const KernelKey &k, const StepArgs<float> &a, int grid, hipStream_t s);

// sf neighbour gather, step_<storage>.hip
hipError_t launch_sf_gather_fp64(int lattice, const StepArgs<double> &a, const int *sfnb, int nsf,
                                 double *out, hipStream_t s);
hipError_t launch_sf_gather_fp32(int lattice, const StepArgs<float> &a, const int *sfnb, int nsf,
                                 float *out, hipStream_t s);
hipError_t launch_sf_gather_fp16s(int lattice, const StepArgs<float> &a, const int *sfnb, int nsf,
                                  float *out, hipStream_t s);
hipError_t launch_sf_gather_bf16s(int lattice, const StepArgs<float> &a, const int *sfnb, int nsf,
                                  float *out, hipStream_t s);

// auxiliary kernels, aux.hip
hipError_t launch_init_static(int lattice, int storage, void *f, long long npad, long long n,
                              hipStream_t s);
// hs: the half-interleaved 16-bit layout's N / 2, or 0 (StepArgs::hs)
hipError_t launch_init_equilibrium(int lattice, int storage, void *f, long long npad, long long n, long long hs,
                                   const double *rho, const double *ux, const double *uy,
                                   const double *uz, hipStream_t s);
// slots first .. first + n - 1 of direction i of f -> double[n]; all n slots <- double[n]
hipError_t launch_f_to_double(int lattice, int storage, const void *f, long long npad, int i,
                              long long first, long long n, long long hs, double *out, hipStream_t s);
hipError_t launch_f_from_double(int lattice, int storage, void *f, long long npad, int i,
                                long long n, long long hs, const double *in, hipStream_t s);
// E partials: one double per block; then the ordered final sum
hipError_t launch_energy(int dim, int cdouble, const void *rho, const void *ux, const void *uy,
                         const void *uz, long long n, double *partials, int nblocks, double *out,
                         hipStream_t s);
hipError_t launch_links(int lattice, const uint8_t *mask, const uint16_t *cls, int Lx, int Ly, int Lz,
                        uint16_t *links, hipStream_t s);
hipError_t launch_copy(const void *src, void *dst, size_t nbytes, hipStream_t s);

// batched step, batch_fp64.hip (the other storages are launched in batch.hip)
hipError_t launch_batch_fp64(int lattice, const void *members, const int2 *chunks, int nchunks, int grid, int sub,
                             int lds, hipStream_t s);

// learned closure, mlp.hip: operand image of a model's current weights with
// the fallback band, into device memory dst on stream s
size_t mlp_image_bytes();
lbmc_status mlp_build_image(const lbmc_mlp *m, void *dst, int device, double lo, double hi, hipStream_t s);
// whether the model is evaluated with fp16 operands (lbmc_mlp_set_eval_fp16)
bool mlp_eval_fp16(const lbmc_mlp *m);

}  // namespace lbmc
