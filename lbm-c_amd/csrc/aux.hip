// Auxiliary device kernels: initial conditions, slot conversion for the
// host read-back, the kinetic-energy reduction, bounceback link sets and the
// bandwidth-reference copy.  Built with contraction off so that the FP64
// equilibrium initial condition rounds operation by operation like the
// reference's FP64 arithmetic.
#include <hip/hip_runtime.h>

#include "dispatch.hpp"
#include "lattice.hpp"
#include "node_ops.hpp"
#include "storage.hpp"

namespace lbmc {

namespace {

constexpr int kAuxBlock = 256;

inline int grid_for(long long n, int per_block = kAuxBlock)
{
    long long g = (n + per_block - 1) / per_block;
    if (g > 1048576) g = 1048576;  // grid-stride loops cover the rest
    return (int)(g < 1 ? 1 : g);
}

template <class L, class St>
__global__ void init_static_kernel(typename St::S *f, long long npad, long long n)
{
    using C = typename St::C;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        sfor<0, L::Q>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            // load_static_IC, LBM-C/model_builder.cu:233-248: f_i = 1.0*w_i
            const double v = St::shifted ? 0.0 : 1.0 * L::w[i];
            f[(size_t)i * npad + k] = St::store((C)v);
        });
    }
}

// element of slot k (step_kernel.hpp slot_elem): the half-interleaved 16-bit
// layout keeps slots k and k + hs in one word
__device__ __forceinline__ long long elem_of(long long k, long long hs)
{
    return hs == 0 ? k : (k < hs ? 2 * k : 2 * (k - hs) + 1);
}

template <class L, class St>
__global__ void init_eq_kernel(typename St::S *f, long long npad, long long n, long long hs, const double *rho,
                               const double *ux, const double *uy, const double *uz)
{
    using C = typename St::C;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        double u[3] = {ux[k], uy[k], L::D > 2 ? uz[k] : 0.0};
        double feq[L::Q];
        equilibrium<L>(rho[k], u, feq);
        sfor<0, L::Q>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            double v = feq[i];
            if constexpr (St::shifted) v = v - L::w[i];
            f[(size_t)i * npad + elem_of(k, hs)] = St::store((C)v);
        });
    }
}

template <class L, class St>
__global__ void f_to_double_kernel(const typename St::S *f, long long npad, int i, long long first, long long n,
                                   long long hs, double *out)
{
    const double w = L::w[i];
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        double v = (double)St::load(f[(size_t)i * npad + elem_of(first + k, hs)]);
        if constexpr (St::shifted) v = w + v;
        out[k] = v;
    }
}

template <class L, class St>
__global__ void f_from_double_kernel(typename St::S *f, long long npad, int i, long long n, long long hs,
                                     const double *in)
{
    using C = typename St::C;
    const double w = L::w[i];
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        double v = in[k];
        if constexpr (St::shifted) v = v - w;
        f[(size_t)i * npad + elem_of(k, hs)] = St::store((C)v);
    }
}

// per-block partial of sum 1/2 rho |u|^2 (the energy functor of
// LBM-C/index.cu:315-335), accumulated in FP64
template <class C, int D>
__global__ void energy_partial_kernel(const C *rho, const C *ux, const C *uy, const C *uz,
                                      long long n, double *partials)
{
    __shared__ double red[kAuxBlock];
    double acc = 0.0;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const double x = ux[k], y = uy[k];
        double q = x * x + y * y;
        if constexpr (D > 2) {
            const double z = uz[k];
            q = q + z * z;
        }
        acc += 0.5 * (double)rho[k] * q;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kAuxBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void energy_final_kernel(const double *partials, int nblocks, double *out)
{
    __shared__ double red[kAuxBlock];
    double acc = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) acc += partials[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kAuxBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <class L>
__global__ void links_kernel(const uint8_t *mask, const uint16_t *cls, int Lx, int Ly, int Lz,
                             uint16_t *links)
{
    const long long n = (long long)Lx * Ly * Lz;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(k % Lx);
        const int y = (int)((k / Lx) % Ly);
        const int z = (int)(k / ((long long)Lx * Ly));
        uint16_t b = 0;
        if (!(cls[mask[k]] & CLS_SOLID)) {
            sfor<1, L::Q>([&](auto ii) {
                constexpr int i = decltype(ii)::value;
                int xn = x + L::ex[i];
                xn = xn < 0 ? xn + Lx : (xn >= Lx ? xn - Lx : xn);
                int yn = y + L::ey[i];
                yn = yn < 0 ? yn + Ly : (yn >= Ly ? yn - Ly : yn);
                int zn = z + L::ez[i];
                zn = zn < 0 ? zn + Lz : (zn >= Lz ? zn - Lz : zn);
                const long long nn = ((long long)zn * Ly + yn) * Lx + xn;
                if (cls[mask[nn]] & CLS_SOLID) b |= (uint16_t)(1u << i);
            });
        }
        links[k] = b;
    }
}

__global__ void copy_kernel(const uint4 *src, uint4 *dst, long long n16)
{
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n16;
         k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

template <class L, class F>
hipError_t by_storage(int storage, F &&fn)
{
    switch (storage) {
    case ST_FP64: return fn(L{}, StFP64{});
    case ST_FP32: return fn(L{}, StFP32{});
    case ST_FP16S: return fn(L{}, StFP16S{});
    case ST_BF16S: return fn(L{}, StBF16S{});
    default: return hipErrorInvalidValue;
    }
}

template <class F>
hipError_t by_lattice_storage(int lattice, int storage, F &&fn)
{
    return lattice == 0 ? by_storage<D2Q9>(storage, fn) : by_storage<D3Q15>(storage, fn);
}

}  // namespace

hipError_t launch_init_static(int lattice, int storage, void *f, long long npad, long long n,
                              hipStream_t s)
{
    return by_lattice_storage(lattice, storage, [&](auto l, auto st) {
        using L = decltype(l);
        using St = decltype(st);
        hipLaunchKernelGGL((init_static_kernel<L, St>), dim3(grid_for(n)), dim3(kAuxBlock), 0, s,
                           (typename St::S *)f, npad, n);
        return hipGetLastError();
    });
}

hipError_t launch_init_equilibrium(int lattice, int storage, void *f, long long npad, long long n, long long hs,
                                   const double *rho, const double *ux, const double *uy,
                                   const double *uz, hipStream_t s)
{
    return by_lattice_storage(lattice, storage, [&](auto l, auto st) {
        using L = decltype(l);
        using St = decltype(st);
        hipLaunchKernelGGL((init_eq_kernel<L, St>), dim3(grid_for(n)), dim3(kAuxBlock), 0, s,
                           (typename St::S *)f, npad, n, hs, rho, ux, uy, uz);
        return hipGetLastError();
    });
}

hipError_t launch_f_to_double(int lattice, int storage, const void *f, long long npad, int i,
                              long long first, long long n, long long hs, double *out, hipStream_t s)
{
    return by_lattice_storage(lattice, storage, [&](auto l, auto st) {
        using L = decltype(l);
        using St = decltype(st);
        hipLaunchKernelGGL((f_to_double_kernel<L, St>), dim3(grid_for(n)), dim3(kAuxBlock), 0, s,
                           (const typename St::S *)f, npad, i, first, n, hs, out);
        return hipGetLastError();
    });
}

hipError_t launch_f_from_double(int lattice, int storage, void *f, long long npad, int i,
                                long long n, long long hs, const double *in, hipStream_t s)
{
    return by_lattice_storage(lattice, storage, [&](auto l, auto st) {
        using L = decltype(l);
        using St = decltype(st);
        hipLaunchKernelGGL((f_from_double_kernel<L, St>), dim3(grid_for(n)), dim3(kAuxBlock), 0, s,
                           (typename St::S *)f, npad, i, n, hs, in);
        return hipGetLastError();
    });
}

hipError_t launch_energy(int dim, int cdouble, const void *rho, const void *ux, const void *uy,
                         const void *uz, long long n, double *partials, int nblocks, double *out,
                         hipStream_t s)
{
    if (cdouble) {
        if (dim > 2)
            hipLaunchKernelGGL((energy_partial_kernel<double, 3>), dim3(nblocks), dim3(kAuxBlock), 0, s,
                               (const double *)rho, (const double *)ux, (const double *)uy,
                               (const double *)uz, n, partials);
        else
            hipLaunchKernelGGL((energy_partial_kernel<double, 2>), dim3(nblocks), dim3(kAuxBlock), 0, s,
                               (const double *)rho, (const double *)ux, (const double *)uy,
                               (const double *)uz, n, partials);
    } else {
        if (dim > 2)
            hipLaunchKernelGGL((energy_partial_kernel<float, 3>), dim3(nblocks), dim3(kAuxBlock), 0, s,
                               (const float *)rho, (const float *)ux, (const float *)uy,
                               (const float *)uz, n, partials);
        else
            hipLaunchKernelGGL((energy_partial_kernel<float, 2>), dim3(nblocks), dim3(kAuxBlock), 0, s,
                               (const float *)rho, (const float *)ux, (const float *)uy,
                               (const float *)uz, n, partials);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(kAuxBlock), 0, s, partials, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_links(int lattice, const uint8_t *mask, const uint16_t *cls, int Lx, int Ly, int Lz,
                        uint16_t *links, hipStream_t s)
{
    const long long n = (long long)Lx * Ly * Lz;
    if (lattice == 0)
        hipLaunchKernelGGL((links_kernel<D2Q9>), dim3(grid_for(n)), dim3(kAuxBlock), 0, s, mask, cls,
                           Lx, Ly, Lz, links);
    else
        hipLaunchKernelGGL((links_kernel<D3Q15>), dim3(grid_for(n)), dim3(kAuxBlock), 0, s, mask, cls,
                           Lx, Ly, Lz, links);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, size_t nbytes, hipStream_t s)
{
    const long long n16 = (long long)(nbytes / 16);
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n16)), dim3(kAuxBlock), 0, s, (const uint4 *)src,
                       (uint4 *)dst, n16);
    return hipGetLastError();
}

}  // namespace lbmc
