// Stand-alone stencil primitives (Laplacian, divergence, gradient), layout
// conversion and material-coefficient preparation.
//
// The primitives back the public laplacian_center / divergence_center /
// gradient_center; the time loop and the projection use their own fused
// kernels. All are HBM-bound row kernels: block (64, 4), lanes along z.
#include "sw_common.h"

// bit-for-bit float32 parity with NumPy: no multiply-add contraction in these kernels
#pragma clang fp contract(off)

namespace sw {
namespace {

template <typename T>
__global__ void __launch_bounds__(256) laplacian_kernel(const T* __restrict__ f, T* __restrict__ out, Grid g,
                                                        T hx2, T hy2, T hz2) {
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    SW_FOR_EACH_VOXEL(g, {
        T v = T(0);
        if (g.interior(x, y, z)) {
            const T c = f[i];
            v = (f[i + P] - T(2) * c + f[i - P]) / hx2;
            v = v + (f[i + R] - T(2) * c + f[i - R]) / hy2;
            v = v + (f[i + 1] - T(2) * c + f[i - 1]) / hz2;
        }
        out[i] = v;
    });
}

template <typename T>
__global__ void __launch_bounds__(256) divergence_kernel(const T* __restrict__ fx, const T* __restrict__ fy,
                                                         const T* __restrict__ fz, T* __restrict__ out, Grid g,
                                                         T h2x, T h2y, T h2z) {
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    SW_FOR_EACH_VOXEL(g, {
        T v = T(0);
        if (g.interior(x, y, z)) {
            v = (fx[i + P] - fx[i - P]) / h2x;
            v = v + (fy[i + R] - fy[i - R]) / h2y;
            v = v + (fz[i + 1] - fz[i - 1]) / h2z;
        }
        out[i] = v;
    });
}

template <typename T>
__global__ void __launch_bounds__(256) gradient_kernel(const T* __restrict__ phi, T* __restrict__ gx,
                                                       T* __restrict__ gy, T* __restrict__ gz, Grid g, T h2x, T h2y,
                                                       T h2z) {
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    SW_FOR_EACH_VOXEL(g, {
        T a = T(0), b = T(0), c = T(0);
        if (g.interior(x, y, z)) {
            a = (phi[i + P] - phi[i - P]) / h2x;
            b = (phi[i + R] - phi[i - R]) / h2y;
            c = (phi[i + 1] - phi[i - 1]) / h2z;
        }
        gx[i] = a;
        gy[i] = b;
        gz[i] = c;
    });
}

template <typename T>
__global__ void __launch_bounds__(256) aos3_to_soa_kernel(const T* __restrict__ aos, T* __restrict__ x,
                                                          T* __restrict__ y, T* __restrict__ z, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] = aos[3 * i];
        y[i] = aos[3 * i + 1];
        z[i] = aos[3 * i + 2];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) soa_to_aos3_kernel(const T* __restrict__ x, const T* __restrict__ y,
                                                          const T* __restrict__ z, T* __restrict__ aos, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        aos[3 * i] = x[i];
        aos[3 * i + 1] = y[i];
        aos[3 * i + 2] = z[i];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) coefficients_kernel(const double* __restrict__ rho, const double* __restrict__ mu,
                                                           const double* __restrict__ eta, T* __restrict__ m,
                                                           T* __restrict__ e, T* __restrict__ ir, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double r = rho[i];
        if (m) m[i] = (T)(mu[i] / r);
        if (e) e[i] = (T)(eta[i] / r);
        if (ir) ir[i] = (T)(1.0 / r);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) import_state_kernel(const T* __restrict__ u0, const T* __restrict__ v0, T dt,
                                                           bool incr, T* __restrict__ ux, T* __restrict__ uy,
                                                           T* __restrict__ uz, T* __restrict__ bx, T* __restrict__ by,
                                                           T* __restrict__ bz, Grid g) {
    SW_FOR_EACH_VOXEL(g, {
        const bool in = g.interior(x, y, z);
        T* U[3] = {ux, uy, uz};
        T* B[3] = {bx, by, bz};
        for (int c = 0; c < 3; ++c) {
            const T u = (in && u0) ? u0[3 * i + c] : T(0);
            // solver.py:164, 170-171: u_prev = mask*(u - dt*v0), or 0 without initial_v
            const T up = (in && v0) ? (u - dt * v0[3 * i + c]) : T(0);
            U[c][i] = u;
            B[c][i] = incr ? (u - up) : up;
        }
    });
}

template <typename T>
__global__ void __launch_bounds__(256) export_state_kernel(const T* __restrict__ ux, const T* __restrict__ uy,
                                                           const T* __restrict__ uz, const T* __restrict__ bx,
                                                           const T* __restrict__ by, const T* __restrict__ bz,
                                                           T inv_dt, bool incr, T* __restrict__ u_aos,
                                                           T* __restrict__ v_aos, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T u[3] = {ux[i], uy[i], uz[i]};
        const T b[3] = {bx[i], by[i], bz[i]};
        for (int c = 0; c < 3; ++c) {
            if (u_aos) u_aos[3 * i + c] = u[c];
            if (v_aos) v_aos[3 * i + c] = (incr ? b[c] : (u[c] - b[c])) * inv_dt;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) combine3_kernel(const T* __restrict__ a0, const T* __restrict__ a1,
                                                       const T* __restrict__ a2, const T* __restrict__ b0,
                                                       const T* __restrict__ b1, const T* __restrict__ b2, T alpha,
                                                       T beta, T* o0, T* o1, T* o2, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T r0 = alpha * a0[i] + beta * b0[i];
        const T r1 = alpha * a1[i] + beta * b1[i];
        const T r2 = alpha * a2[i] + beta * b2[i];
        o0[i] = r0;
        o1[i] = r1;
        o2[i] = r2;
    }
}

// out[k] = f[idx[k]]; an index outside [0, n) reads nothing and yields NaN.
template <typename T>
__global__ void __launch_bounds__(256) gather_kernel(const T* __restrict__ f, int64_t n,
                                                     const int64_t* __restrict__ idx, int64_t count,
                                                     T* __restrict__ out, int64_t stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        const int64_t j = idx[k];
        out[k * stride] = (j >= 0 && j < n) ? f[j] : (T)NAN;
    }
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int sw_laplacian_center(int dtype, const void* f, void* out, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                        double dz, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(f && out, SW_ERR_ARG, "null pointer");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(laplacian_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0, (hipStream_t)stream,
                                   (const T*)f, (T*)out, g, (T)(dx * dx), (T)(dy * dy), (T)(dz * dz)));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_divergence_center(int dtype, const void* fx, const void* fy, const void* fz, void* out, int64_t nx, int64_t ny,
                         int64_t nz, double dx, double dy, double dz, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(fx && fy && fz && out, SW_ERR_ARG, "null pointer");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(divergence_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0, (hipStream_t)stream,
                                   (const T*)fx, (const T*)fy, (const T*)fz, (T*)out, g, (T)(2.0 * dx),
                                   (T)(2.0 * dy), (T)(2.0 * dz)));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_gradient_center(int dtype, const void* phi, void* gx, void* gy, void* gz, int64_t nx, int64_t ny, int64_t nz,
                       double dx, double dy, double dz, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(phi && gx && gy && gz, SW_ERR_ARG, "null pointer");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(gradient_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0, (hipStream_t)stream,
                                   (const T*)phi, (T*)gx, (T*)gy, (T*)gz, g, (T)(2.0 * dx), (T)(2.0 * dy),
                                   (T)(2.0 * dz)));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_aos3_to_soa(int dtype, const void* aos, void* x, void* y, void* z, int64_t count, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(count >= 0, SW_ERR_ARG, "negative count");
    if (count == 0) return SW_OK;
    SW_REQUIRE(aos && x && y && z, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(aos3_to_soa_kernel<T>, dim3(flat_grid(count)), dim3(256), 0, (hipStream_t)stream,
                                   (const T*)aos, (T*)x, (T*)y, (T*)z, count));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_soa_to_aos3(int dtype, const void* x, const void* y, const void* z, void* aos, int64_t count, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(count >= 0, SW_ERR_ARG, "negative count");
    if (count == 0) return SW_OK;
    SW_REQUIRE(aos && x && y && z, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(soa_to_aos3_kernel<T>, dim3(flat_grid(count)), dim3(256), 0, (hipStream_t)stream,
                                   (const T*)x, (const T*)y, (const T*)z, (T*)aos, count));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_prepare_coefficients(int dtype, const double* rho, const double* mu, const double* eta, void* mu_over_rho,
                            void* eta_over_rho, void* inv_rho, int64_t n, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(n >= 0, SW_ERR_ARG, "negative count");
    if (n == 0) return SW_OK;
    SW_REQUIRE(rho, SW_ERR_ARG, "rho is null");
    SW_REQUIRE(!mu_over_rho || mu, SW_ERR_ARG, "mu is null");
    SW_REQUIRE(!eta_over_rho || eta, SW_ERR_ARG, "eta is null but eta_over_rho requested");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(coefficients_kernel<T>, dim3(flat_grid(n)), dim3(256), 0, (hipStream_t)stream, rho,
                                   mu, eta, (T*)mu_over_rho, (T*)eta_over_rho, (T*)inv_rho, n));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_import_state(int dtype, int form, const void* u0, const void* v0, double dt, void* const* u, void* const* b,
                    int64_t nx, int64_t ny, int64_t nz, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(form == SW_FORM_INCREMENT || form == SW_FORM_LEAPFROG, SW_ERR_ARG, "unknown form %d", form);
    SW_REQUIRE(u && b && u[0] && u[1] && u[2] && b[0] && b[1] && b[2], SW_ERR_ARG, "null state buffer");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(import_state_kernel<T>, dim3(row_grid(g, 0)), row_block(), 0, (hipStream_t)stream,
                                   (const T*)u0, (const T*)v0, (T)dt, form == SW_FORM_INCREMENT, (T*)u[0], (T*)u[1],
                                   (T*)u[2], (T*)b[0], (T*)b[1], (T*)b[2], g));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_export_state(int dtype, int form, const void* const* u, const void* const* b, double inv_dt, void* u_aos,
                    void* v_aos, int64_t n, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(n >= 0, SW_ERR_ARG, "negative count");
    if (n == 0) return SW_OK;
    SW_REQUIRE(form == SW_FORM_INCREMENT || form == SW_FORM_LEAPFROG, SW_ERR_ARG, "unknown form %d", form);
    SW_REQUIRE(u && b && u[0] && u[1] && u[2] && b[0] && b[1] && b[2], SW_ERR_ARG, "null state buffer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(export_state_kernel<T>, dim3(flat_grid(n)), dim3(256), 0, (hipStream_t)stream,
                                   (const T*)u[0], (const T*)u[1], (const T*)u[2], (const T*)b[0], (const T*)b[1],
                                   (const T*)b[2], (T)inv_dt, form == SW_FORM_INCREMENT, (T*)u_aos, (T*)v_aos, n));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_combine3(int dtype, const void* const* a, const void* const* b, double alpha, double beta, void* const* out,
                int64_t n, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(n >= 0, SW_ERR_ARG, "negative count");
    if (n == 0) return SW_OK;
    SW_REQUIRE(a && b && out, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(combine3_kernel<T>, dim3(flat_grid(n)), dim3(256), 0, (hipStream_t)stream,
                                   (const T*)a[0], (const T*)a[1], (const T*)a[2], (const T*)b[0], (const T*)b[1],
                                   (const T*)b[2], (T)alpha, (T)beta, (T*)out[0], (T*)out[1], (T*)out[2], n));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_gather_traces(int dtype, const void* f, int64_t n, const int64_t* idx, int64_t count, void* out,
                     int64_t out_stride, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(count >= 0 && out_stride >= 1 && n >= 0, SW_ERR_ARG, "bad count/stride/n");
    if (count == 0) return SW_OK;
    SW_REQUIRE(f && idx && out, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(gather_kernel<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                                   (hipStream_t)stream, (const T*)f, n, idx, count, (T*)out, out_stride));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
