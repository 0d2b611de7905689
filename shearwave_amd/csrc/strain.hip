// Strain post-processing on gfx950: the Voigt strain tensor from nine 2h
// central partials (strain.py:16-70) and its four scalar invariants
// (strain.py:73-135), separately (drop-in) and fused (displacement in,
// invariants out, no tensor in memory: 12 + 16 B per voxel in float32).
#include "sw_common.h"

// bit-for-bit float32 parity with NumPy: no multiply-add contraction in these kernels
#pragma clang fp contract(off)

namespace sw {
namespace {

// Nine partials of (ux, uy, uz) at voxel i; zero on the boundary layer.
// Arithmetic as gradient_center (solver.py:720-729): (f[+1] - f[-1]) / (2h) in T.
template <typename T>
__device__ inline void voigt_at(const T* __restrict__ ux, const T* __restrict__ uy, const T* __restrict__ uz,
                                const Grid& g, int x, int y, int z, int64_t i, T h2x, T h2y, T h2z, T* e) {
    if (!g.interior(x, y, z)) {
#pragma unroll
        for (int k = 0; k < 6; ++k) e[k] = T(0);
        return;
    }
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    const T ux_x = (ux[i + P] - ux[i - P]) / h2x, ux_y = (ux[i + R] - ux[i - R]) / h2y, ux_z = (ux[i + 1] - ux[i - 1]) / h2z;
    const T uy_x = (uy[i + P] - uy[i - P]) / h2x, uy_y = (uy[i + R] - uy[i - R]) / h2y, uy_z = (uy[i + 1] - uy[i - 1]) / h2z;
    const T uz_x = (uz[i + P] - uz[i - P]) / h2x, uz_y = (uz[i + R] - uz[i - R]) / h2y, uz_z = (uz[i + 1] - uz[i - 1]) / h2z;
    e[0] = ux_x;
    e[1] = uy_y;
    e[2] = uz_z;
    e[3] = uy_z + uz_y;
    e[4] = ux_z + uz_x;
    e[5] = ux_y + uy_x;
}

// Invariants in the reference's order of operations (strain.py:110-128), in T;
// von Mises is also returned in double as sqrt(2/3) * (double)dev.
template <typename T>
__device__ inline void invariants(const T* e, T& vol, T& dev, T& ms, double& vm) {
    vol = ((e[0] + e[1]) + e[2]) / T(3);
    const T dxx = e[0] - vol, dyy = e[1] - vol, dzz = e[2] - vol;
    const T nrm = (dxx * dxx + dyy * dyy) + dzz * dzz;
    const T shr = (e[3] * e[3] + e[4] * e[4]) + e[5] * e[5];
    dev = sqrt(nrm + T(0.5) * shr);
    const T j2 = T(0.5) * (nrm + T(0.5) * shr);
    ms = sqrt(j2 > T(0) ? j2 : T(0));
    vm = 0.816496580927726 * (double)dev;  // sqrt(2/3)
}

template <typename T>
__global__ void __launch_bounds__(256) strain_tensor_kernel(const T* __restrict__ ux, const T* __restrict__ uy,
                                                            const T* __restrict__ uz, T* __restrict__ eps, Grid g,
                                                            T h2x, T h2y, T h2z) {
    SW_FOR_EACH_VOXEL(g, {
        T e[6];
        voigt_at(ux, uy, uz, g, x, y, z, i, h2x, h2y, h2z, e);
        for (int k = 0; k < 6; ++k) eps[6 * i + k] = e[k];
    });
}

template <typename T>
__global__ void __launch_bounds__(256) strain_invariants_kernel(const T* __restrict__ eps, int64_t n,
                                                                T* __restrict__ vol, T* __restrict__ dev,
                                                                T* __restrict__ ms, double* __restrict__ vm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T e[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) e[k] = eps[6 * i + k];
        T a, b, c;
        double d;
        invariants(e, a, b, c, d);
        vol[i] = a;
        dev[i] = b;
        ms[i] = c;
        vm[i] = d;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) strain_invariants_from_u_kernel(const T* __restrict__ ux,
                                                                       const T* __restrict__ uy,
                                                                       const T* __restrict__ uz, T* __restrict__ vol,
                                                                       T* __restrict__ dev, T* __restrict__ ms,
                                                                       T* __restrict__ vm, Grid g, T h2x, T h2y,
                                                                       T h2z) {
    SW_FOR_EACH_VOXEL(g, {
        T e[6];
        voigt_at(ux, uy, uz, g, x, y, z, i, h2x, h2y, h2z, e);
        T a, b, c;
        double d;
        invariants(e, a, b, c, d);
        vol[i] = a;
        dev[i] = b;
        ms[i] = c;
        vm[i] = (T)d;
    });
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int sw_strain_tensor(int dtype, const void* ux, const void* uy, const void* uz, void* eps, int64_t nx, int64_t ny,
                     int64_t nz, double dx, double dy, double dz, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(ux && uy && uz && eps, SW_ERR_ARG, "null pointer");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(strain_tensor_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0, (hipStream_t)stream,
                                   (const T*)ux, (const T*)uy, (const T*)uz, (T*)eps, g, (T)(2.0 * dx), (T)(2.0 * dy),
                                   (T)(2.0 * dz)));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_strain_invariants(int dtype, const void* eps, int64_t count, void* vol, void* dev, void* max_shear,
                         void* von_mises, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(count >= 0, SW_ERR_ARG, "negative count");
    if (count == 0) return SW_OK;
    SW_REQUIRE(eps && vol && dev && max_shear && von_mises, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(strain_invariants_kernel<T>, dim3(flat_grid(count)), dim3(256), 0,
                                   (hipStream_t)stream, (const T*)eps, count, (T*)vol, (T*)dev, (T*)max_shear,
                                   (double*)von_mises));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_strain_invariants_from_u(int dtype, const void* ux, const void* uy, const void* uz, void* vol, void* dev,
                                void* max_shear, void* von_mises, int64_t nx, int64_t ny, int64_t nz, double dx,
                                double dy, double dz, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(ux && uy && uz && vol && dev && max_shear && von_mises, SW_ERR_ARG, "null pointer");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(strain_invariants_from_u_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0,
                                   (hipStream_t)stream, (const T*)ux, (const T*)uy, (const T*)uz, (T*)vol, (T*)dev,
                                   (T*)max_shear, (T*)von_mises, g, (T)(2.0 * dx), (T)(2.0 * dy), (T)(2.0 * dz)));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
