// Adjoint of the fused time step, for gradients of a displacement misfit with
// respect to the mu and eta fields (elastography inversion). The reference has
// no derivative code; this is the exact transpose of the update the step
// kernels apply (increment form, solver.py:201-226 restated):
//
//   forward, interior i:  d'_i = M_i g (d_i + m_i (L u)_i + e_i (Lv d)_i + F_i),  u'_i = u_i + d'_i
//   with (L u)_i = sum_a dt^2/h_a^2 (u_{i+a} + u_{i-a} - 2 u_i), (Lv d)_i likewise with dt/h_a^2,
//   M the interior mask and g = 1 - vel_damp.
//
//   adjoint, given (ub', db') of (u', d'):  w = M g (ub' + db')
//     ub = ub' + L (m w)          db = w + Lv (e w)
//     dJ/dm_i += sum_c w_ic (L u_c)_i      dJ/de_i += sum_c w_ic (Lv d_c)_i
//
// L is symmetric on the interior (w vanishes on the boundary layer), so L^T
// is L applied to the products. The two gradient sums accumulate in float64.
#include "sw_common.h"

namespace sw {
namespace {

template <typename T>
struct AdjArgs {
    const T* u[3];   // forward state u_n
    const T* d[3];   // forward state d_n
    const T* ub1[3]; // adjoint of u_{n+1}
    const T* db1[3]; // adjoint of d_{n+1}
    T* ub[3];        // adjoint of u_n (out)
    T* db[3];        // adjoint of d_n (out)
    const T* m;
    const T* e;      // NULL when inviscid
    double* gm;      // dJ/d(mu/rho), accumulated
    double* ge;      // dJ/d(eta/rho), accumulated (ignored when inviscid)
    Grid g;
    T cx, cy, cz, vx, vy, vz, damp;
};

template <typename T, bool VISC>
__global__ void __launch_bounds__(256) adjoint_step_kernel(AdjArgs<T> a) {
    const Grid g = a.g;
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    SW_FOR_EACH_VOXEL(g, {
        T oub[3] = {T(0), T(0), T(0)}, odb[3] = {T(0), T(0), T(0)};
        if (g.interior(x, y, z)) {
            double gmi = 0.0, gei = 0.0;
            const T mi = a.m[i];
            const T ei = VISC ? a.e[i] : T(0);
            // w at a voxel: zero on the boundary layer
            auto inside = [&](int xx, int yy, int zz) { return g.interior(xx, yy, zz); };
            const bool xp = inside(x + 1, y, z), xm = inside(x - 1, y, z);
            const bool yp = inside(x, y + 1, z), ym = inside(x, y - 1, z);
            const bool zp = inside(x, y, z + 1), zm = inside(x, y, z - 1);
            for (int c = 0; c < 3; ++c) {
                auto w = [&](int64_t k, bool in) { return in ? a.damp * (a.ub1[c][i + k] + a.db1[c][i + k]) : T(0); };
                const T w0 = w(0, true);
                const T wxp = w(P, xp), wxm = w(-P, xm), wyp = w(R, yp), wym = w(-R, ym), wzp = w(1, zp),
                        wzm = w(-1, zm);
                const T mw0 = mi * w0;
                const T lmw = a.cx * (a.m[i + P] * wxp + a.m[i - P] * wxm - T(2) * mw0) +
                              a.cy * (a.m[i + R] * wyp + a.m[i - R] * wym - T(2) * mw0) +
                              a.cz * (a.m[i + 1] * wzp + a.m[i - 1] * wzm - T(2) * mw0);
                oub[c] = a.ub1[c][i] + lmw;
                T dbv = w0;
                const T* U = a.u[c];
                const T lu = a.cx * (U[i + P] + U[i - P] - T(2) * U[i]) + a.cy * (U[i + R] + U[i - R] - T(2) * U[i]) +
                             a.cz * (U[i + 1] + U[i - 1] - T(2) * U[i]);
                gmi += (double)w0 * (double)lu;
                if (VISC) {
                    const T ew0 = ei * w0;
                    const T lew = a.vx * (a.e[i + P] * wxp + a.e[i - P] * wxm - T(2) * ew0) +
                                  a.vy * (a.e[i + R] * wyp + a.e[i - R] * wym - T(2) * ew0) +
                                  a.vz * (a.e[i + 1] * wzp + a.e[i - 1] * wzm - T(2) * ew0);
                    dbv += lew;
                    const T* D = a.d[c];
                    const T ld = a.vx * (D[i + P] + D[i - P] - T(2) * D[i]) +
                                 a.vy * (D[i + R] + D[i - R] - T(2) * D[i]) +
                                 a.vz * (D[i + 1] + D[i - 1] - T(2) * D[i]);
                    gei += (double)w0 * (double)ld;
                }
                odb[c] = dbv;
            }
            a.gm[i] += gmi;
            if (VISC) a.ge[i] += gei;
        }
        for (int c = 0; c < 3; ++c) {
            a.ub[c][i] = oub[c];
            a.db[c][i] = odb[c];
        }
    });
}

template <typename T>
__global__ void __launch_bounds__(256) inject_kernel(T* __restrict__ f, const int64_t* __restrict__ idx, int64_t count,
                                                     const T* __restrict__ grad, int64_t stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) f[idx[k]] += grad[k * stride];
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int sw_adjoint_step(int dtype, const void* const* u, const void* const* d, const void* const* ub_next,
                    const void* const* db_next, void* const* ub, void* const* db, const void* mu_over_rho,
                    const void* eta_over_rho, double* grad_m, double* grad_e, int64_t nx, int64_t ny, int64_t nz,
                    double dx, double dy, double dz, double dt, double vel_damp, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(u && d && ub_next && db_next && ub && db && mu_over_rho && grad_m, SW_ERR_ARG, "null pointer");
    SW_REQUIRE(!eta_over_rho || grad_e, SW_ERR_ARG, "grad_e is null");
    SW_REQUIRE(dt > 0 && dx > 0 && dy > 0 && dz > 0, SW_ERR_ARG, "dt and spacing must be > 0");
    Grid g{(int)nx, (int)ny, (int)nz};
    SW_DISPATCH(dtype, T, {
        AdjArgs<T> a{};
        for (int c = 0; c < 3; ++c) {
            a.u[c] = (const T*)u[c];
            a.d[c] = (const T*)d[c];
            a.ub1[c] = (const T*)ub_next[c];
            a.db1[c] = (const T*)db_next[c];
            a.ub[c] = (T*)ub[c];
            a.db[c] = (T*)db[c];
        }
        a.m = (const T*)mu_over_rho;
        a.e = (const T*)eta_over_rho;
        a.gm = grad_m;
        a.ge = grad_e;
        a.g = g;
        a.cx = (T)(dt * dt / (dx * dx));
        a.cy = (T)(dt * dt / (dy * dy));
        a.cz = (T)(dt * dt / (dz * dz));
        a.vx = (T)(dt / (dx * dx));
        a.vy = (T)(dt / (dy * dy));
        a.vz = (T)(dt / (dz * dz));
        a.damp = (T)(1.0 - vel_damp);
        if (eta_over_rho)
            hipLaunchKernelGGL((adjoint_step_kernel<T, true>), dim3(row_grid(g, 0)), row_block(), 0,
                               (hipStream_t)stream, a);
        else
            hipLaunchKernelGGL((adjoint_step_kernel<T, false>), dim3(row_grid(g, 0)), row_block(), 0,
                               (hipStream_t)stream, a);
    });
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_inject_traces(int dtype, void* f, const int64_t* idx, int64_t count, const void* grad, int64_t stride,
                     void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(count >= 0 && stride >= 1, SW_ERR_ARG, "bad count/stride");
    if (count == 0) return SW_OK;
    SW_REQUIRE(f && idx && grad, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(inject_kernel<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                                   (hipStream_t)stream, (T*)f, idx, count, (const T*)grad, stride));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
