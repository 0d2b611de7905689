// Strain post-processing on gfx950: the Voigt strain tensor from nine 2h
// central partials (strain.py:16-70) and its four scalar invariants
// (strain.py:73-135), separately (drop-in) and fused (displacement in,
// invariants out, no tensor in memory: 12 + 16 B per voxel in float32).
#include <type_traits>

#include "strain_point.h"
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
        strain_invariants_of(e, a, b, c, d);
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
        strain_invariants_of(e, a, b, c, d);
        vol[i] = a;
        dev[i] = b;
        ms[i] = c;
        vm[i] = (T)d;
    });
}

// ----------------------------------------------------------------------------
// The in-loop path: invariants of u every few steps of sw_fdtd_run. A (BY x 64)
// tile of (y, z) columns marches along x with u at planes x-1 .. x+2 of its
// column in registers and plane x staged in LDS with a one-voxel y/z halo
// (double-buffered: one barrier per plane), so u is read from HBM about once
// (plus 2 halo rows per BY) and the four invariants are written once. The
// partials multiply by 1/(2h) instead of dividing by 2h, one rounding apart
// from the drop-in kernels above (which reproduce NumPy bit for bit).
#pragma clang fp contract(on)

template <typename T>
struct U3 {
    T a, b, c;
};

constexpr int kInvBY = 16;  // tile rows (block of 64 x 16 threads)

template <typename T, int BY>
__global__ void __launch_bounds__(64 * BY) strain_inv_march_kernel(const T* __restrict__ ux, const T* __restrict__ uy,
                                                                   const T* __restrict__ uz, T* __restrict__ vol,
                                                                   T* __restrict__ dev, T* __restrict__ ms,
                                                                   T* __restrict__ vm, Grid g, T ihx, T ihy, T ihz,
                                                                   int xchunk, int nty, int ntz) {
    __shared__ T lds[2][3][BY + 2][66];
    const unsigned nb = gridDim.x;
    unsigned blk = blockIdx.x;
    if ((nb & 7u) == 0u) blk = (blk & 7u) * (nb >> 3) + (blk >> 3);  // XCD-aware: neighbouring tiles share an L2
    const int tzi = (int)(blk % (unsigned)ntz);
    const unsigned rest = blk / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int x0 = (int)(rest / (unsigned)nty) * xchunk;
    const int x1 = min(x0 + xchunk, g.nx);
    const int y0 = tyi * BY, z0 = tzi * 64;
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int y = y0 + ty, z = z0 + tz;
    const bool inyz = y < g.ny && z < g.nz;
    const bool yz_interior = inyz && y > 0 && y < g.ny - 1 && z > 0 && z < g.nz - 1;
    const uint32_t yz = inyz ? (uint32_t)(y * g.nz + z) * (uint32_t)sizeof(T) : 0u;
    // halo: rows y0-1 and y0+BY by the first and last row of threads, columns
    // z0-1 and z0+64 by the first and last lane; points outside the grid are
    // read at offset 0 and only feed boundary voxels, whose invariants are 0
    const bool row_thread = ty == 0 || ty == BY - 1;
    const int ry = ty == 0 ? y0 - 1 : y0 + BY, rly = ty == 0 ? 0 : BY + 1;
    const uint32_t ryz = (row_thread && ry >= 0 && ry < g.ny && z < g.nz) ? (uint32_t)(ry * g.nz + z) * sizeof(T) : 0u;
    const bool col_thread = tz == 0 || tz == 63;
    const int cz = tz == 0 ? z0 - 1 : z0 + 64, clz = tz == 0 ? 0 : 65;
    const uint32_t cyz = (col_thread && cz >= 0 && cz < g.nz && y < g.ny) ? (uint32_t)(y * g.nz + cz) * sizeof(T) : 0u;
    const int64_t P = g.plane();
    auto po = [&](int xx) { return (int64_t)min(max(xx, 0), g.nx - 1) * P; };
    auto ld = [&](int xx, uint32_t o) {
        const int64_t b = po(xx);
        return U3<T>{at_byte(ux + b, o), at_byte(uy + b, o), at_byte(uz + b, o)};
    };
    auto center = [&](int xx) { return ld(xx, yz); };
    auto row_halo = [&](int xx) { return row_thread ? ld(xx, ryz) : U3<T>{}; };
    auto col_halo = [&](int xx) { return col_thread ? ld(xx, cyz) : U3<T>{}; };

    auto plane = [&](auto buf, const U3<T>& m, const U3<T>& c, const U3<T>& p, U3<T>& n, const U3<T>& hc, U3<T>& hn,
                     const U3<T>& vc, U3<T>& vn, int x) {
        constexpr int B = decltype(buf)::value;
        auto& L = lds[B];
        L[0][ty + 1][tz + 1] = c.a;
        L[1][ty + 1][tz + 1] = c.b;
        L[2][ty + 1][tz + 1] = c.c;
        if (row_thread) {
            L[0][rly][tz + 1] = hc.a;
            L[1][rly][tz + 1] = hc.b;
            L[2][rly][tz + 1] = hc.c;
        }
        if (col_thread) {
            L[0][ty + 1][clz] = vc.a;
            L[1][ty + 1][clz] = vc.b;
            L[2][ty + 1][clz] = vc.c;
        }
        n = center(x + 2);
        hn = row_halo(x + 1);
        vn = col_halo(x + 1);
        __syncthreads();
        if (!inyz) return;
        const int r = ty + 1, q = tz + 1;
        auto dy = [&](int f) { return L[f][r + 1][q] - L[f][r - 1][q]; };
        auto dz = [&](int f) { return L[f][r][q + 1] - L[f][r][q - 1]; };
        const UDiff<T> d{p.a - m.a, p.b - m.b, p.c - m.c, dy(0), dy(1), dy(2), dz(0), dz(1), dz(2)};
        T a, b, sh, v;
        inloop_invariants(yz_interior && x > 0 && x < g.nx - 1, d, ihx, ihy, ihz, a, b, sh, v);
        const int64_t px = (int64_t)x * P;
        store_out(&at_byte(vol + px, yz), a);
        store_out(&at_byte(dev + px, yz), b);
        store_out(&at_byte(ms + px, yz), sh);
        store_out(&at_byte(vm + px, yz), v);
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    U3<T> q0 = center(x0 - 1), q1 = center(x0), q2 = center(x0 + 1), q3;
    U3<T> h0 = row_halo(x0), h1, c0 = col_halo(x0), c1;
    // unrolled by four so the queue's roles rotate without register moves
    for (int x = x0; x < x1; x += 4) {
        plane(B0{}, q0, q1, q2, q3, h0, h1, c0, c1, x);
        if (x + 1 >= x1) break;
        plane(B1{}, q1, q2, q3, q0, h1, h0, c1, c0, x + 1);
        if (x + 2 >= x1) break;
        plane(B0{}, q2, q3, q0, q1, h0, h1, c0, c1, x + 2);
        if (x + 3 >= x1) break;
        plane(B1{}, q3, q0, q1, q2, h1, h0, c1, c0, x + 3);
    }
}

#pragma clang fp contract(off)

template <typename T>
void launch_inv_march(const T* ux, const T* uy, const T* uz, T* out, const Grid& g, double dx, double dy, double dz,
                      hipStream_t s) {
    constexpr int BY = kInvBY;
    const int nty = (g.ny + BY - 1) / BY, ntz = (g.nz + 63) / 64;
    // about two blocks per CU in one wave, chunks of at least 16 planes
    const int64_t cols = (int64_t)nty * ntz, target = (int64_t)device_cu_count() * 2;
    int64_t ntx = (target + cols - 1) / cols;
    if (ntx < 1) ntx = 1;
    int xc = (int)((g.nx + ntx - 1) / ntx);
    if (xc < 16) xc = 16;
    ntx = (g.nx + xc - 1) / xc;
    const int64_t n = g.size();
    hipLaunchKernelGGL((strain_inv_march_kernel<T, BY>), dim3((unsigned)(cols * ntx)), dim3(64, BY), 0, s, ux, uy, uz,
                       out, out + n, out + 2 * n, out + 3 * n, g, (T)(0.5 / dx), (T)(0.5 / dy), (T)(0.5 / dz), xc,
                       nty, ntz);
}

}  // namespace

int strain_invariants_in_loop(int dtype, const void* ux, const void* uy, const void* uz, void* out, int64_t nx,
                              int64_t ny, int64_t nz, double dx, double dy, double dz, hipStream_t s) {
    const Grid g{(int)nx, (int)ny, (int)nz};
    if (nz < 32 || ny < 8 || g.plane() >= kMaxMarchPlane) {  // small grids: the per-voxel kernel
        const int64_t n = g.size();
        SW_DISPATCH(dtype, T,
                    hipLaunchKernelGGL(strain_invariants_from_u_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0, s,
                                       (const T*)ux, (const T*)uy, (const T*)uz, (T*)out, (T*)out + n,
                                       (T*)out + 2 * n, (T*)out + 3 * n, g, (T)(2.0 * dx), (T)(2.0 * dy),
                                       (T)(2.0 * dz)));
    } else {
        SW_DISPATCH(dtype, T,
                    launch_inv_march<T>((const T*)ux, (const T*)uy, (const T*)uz, (T*)out, g, dx, dy, dz, s));
    }
    SW_LAUNCH_CHECK();
    return SW_OK;
}

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
