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
#include <cstdlib>

#include "march2_step.h"
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

// ---------------------------------------------------------------------------
// x-marching adjoint tile (grids the march tiles fit). A (BY x 64) tile of
// (y, z) columns marches along x. A point's loads (ub', db', u, d, m, e) are
// issued once, two planes ahead for the tile's own columns and one plane ahead
// for the halo points, and turned into the fields the stencils need:
// m w and e w (w = M g (ub' + db'), zero off the interior), u and d. Plane x of
// those is staged in LDS with a one-voxel halo (rows y0-1 and y0+BY loaded by
// the first and last row of threads, columns z0-1 and z0+64 by the first and
// last lane of each row); the x neighbours come from the thread's own queue.
// Same products and sums as the per-voxel kernel above; the gradient sums
// accumulate in float64.

// Tile rows: 8 in float32 (181 VGPRs for the queues); 4 in float64, whose
// queues hold twice the registers (the compiler keeps the excess in AGPRs,
// one wave per SIMD, no scratch; 8 rows would spill to scratch).
template <typename T>
constexpr int adj_by() {
    return sizeof(T) == 4 ? 8 : 4;
}

template <typename T, bool VISC>
struct AdjRaw {
    T ub[3], db[3], u[3], d[3], m, e;
};

// Stencil fields of a point (w and ub' are kept for the tile's own columns).
template <typename T>
struct AdjPt {
    T mw[3], ew[3], u[3], d[3], w[3], ub[3];
};

template <typename T, bool VISC>
__device__ __forceinline__ AdjRaw<T, VISC> adj_load(const AdjArgs<T>& a, int64_t po, uint32_t o) {
    AdjRaw<T, VISC> r;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.ub[c] = at_byte(a.ub1[c] + po, o);
        r.db[c] = at_byte(a.db1[c] + po, o);
        r.u[c] = at_byte(a.u[c] + po, o);
        r.d[c] = VISC ? (T)at_byte(a.d[c] + po, o) : T(0);
    }
    r.m = at_byte(a.m + po, o);
    r.e = VISC ? (T)at_byte(a.e + po, o) : T(0);
    return r;
}

template <typename T, bool VISC>
__device__ __forceinline__ AdjPt<T> adj_point(const AdjArgs<T>& a, const AdjRaw<T, VISC>& r, bool in) {
    AdjPt<T> p;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T w = in ? a.damp * (r.ub[c] + r.db[c]) : T(0);
        p.w[c] = w;
        p.ub[c] = r.ub[c];
        p.mw[c] = r.m * w;
        p.ew[c] = VISC ? r.e * w : T(0);
        p.u[c] = r.u[c];
        p.d[c] = r.d[c];
    }
    return p;
}

template <typename T, bool VISC, int BY>
struct AdjLds {
    static constexpr int NF = VISC ? 12 : 6;  // m w, u and, with viscosity, e w, d (3 components each)
    T f[2][NF][BY + 2][66];
};

template <typename T, bool VISC, typename Plane>
__device__ __forceinline__ void adj_stage(Plane& L, int r, int q, const AdjPt<T>& p) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        L[c][r][q] = p.mw[c];
        L[3 + c][r][q] = p.u[c];
        if (VISC) {
            L[6 + c][r][q] = p.ew[c];
            L[9 + c][r][q] = p.d[c];
        }
    }
}

template <typename T, bool VISC, int BY>
__global__ void __launch_bounds__(64 * BY) adjoint_march_kernel(AdjArgs<T> a, int xchunk, int nty, int ntz) {
    __shared__ AdjLds<T, VISC, BY> lds;
    const Grid g = a.g;
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);  // XCD-aware: neighbouring tiles share an L2
    const int tzi = (int)(b % (unsigned)ntz);
    const unsigned rest = b / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int x0 = (int)(rest / (unsigned)nty) * xchunk, x1 = min(x0 + xchunk, g.nx);
    const int y0 = tyi * BY, z0 = tzi * 64;
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int y = y0 + ty, z = z0 + tz;
    const bool inyz = y < g.ny && z < g.nz;
    const int64_t P = g.plane();
    const uint32_t yz_el = inyz ? (uint32_t)(y * g.nz + z) : 0u;
    const uint32_t yz = yz_el * (uint32_t)sizeof(T);
    // halo points: a row halo for the first and last row of threads, a column halo for the first and last lane
    const bool row_thread = ty == 0 || ty == BY - 1;
    const int ry = ty == 0 ? y0 - 1 : y0 + BY, rly = ty == 0 ? 0 : BY + 1;
    const bool row_in = row_thread && ry >= 0 && ry < g.ny && z < g.nz;
    const uint32_t ryz = row_in ? (uint32_t)(ry * g.nz + z) * (uint32_t)sizeof(T) : 0u;
    const bool col_thread = tz == 0 || tz == 63;
    const int cz = tz == 0 ? z0 - 1 : z0 + 64, clz = tz == 0 ? 0 : 65;
    const bool col_in = col_thread && cz >= 0 && cz < g.nz && y < g.ny;
    const uint32_t cyz = col_in ? (uint32_t)(y * g.nz + cz) * (uint32_t)sizeof(T) : 0u;

    auto po = [&](int xx) { return (int64_t)min(max(xx, 0), g.nx - 1) * P; };
    auto own = [&](int xx) { return adj_load<T, VISC>(a, po(xx), yz); };
    auto row_raw = [&](int xx) {
        AdjRaw<T, VISC> r{};
        if (row_thread) r = adj_load<T, VISC>(a, po(xx), ryz);
        return r;
    };
    auto col_raw = [&](int xx) {
        AdjRaw<T, VISC> r{};
        if (col_thread) r = adj_load<T, VISC>(a, po(xx), cyz);
        return r;
    };
    auto own_pt = [&](const AdjRaw<T, VISC>& r, int xx) { return adj_point<T, VISC>(a, r, inyz && g.interior(xx, y, z)); };

    // rings indexed by plane: slot (plane - x0) & 3
    AdjPt<T> pt[4];        // own column, planes x-1 .. x+1
    AdjRaw<T, VISC> rq[4]; // own column raw loads, planes x+1 .. x+3
    AdjRaw<T, VISC> hr[4]; // row halo raw loads, planes x .. x+1
    AdjRaw<T, VISC> hc[4]; // column halo raw loads, planes x .. x+1
    pt[3] = own_pt(own(x0 - 1), x0 - 1);
    pt[0] = own_pt(own(x0), x0);
    rq[1] = own(x0 + 1);
    rq[2] = own(x0 + 2);
    hr[0] = row_raw(x0);
    hc[0] = col_raw(x0);

    march_incl<4>(x0, x1 - 1, [&](auto kc, int x) {
        constexpr int K = decltype(kc)::value;
        constexpr int B = K & 1;
        auto& L = lds.f[B];
        // plane x+1 of the own column enters the window; loads run two planes ahead
        pt[(K + 1) & 3] = own_pt(rq[(K + 1) & 3], x + 1);
        rq[(K + 3) & 3] = own(x + 3);
        adj_stage<T, VISC>(L, ty + 1, tz + 1, pt[K]);
        if (row_thread)
            adj_stage<T, VISC>(L, rly, tz + 1, adj_point<T, VISC>(a, hr[K], row_in && g.interior(x, ry, z)));
        if (col_thread)
            adj_stage<T, VISC>(L, ty + 1, clz, adj_point<T, VISC>(a, hc[K], col_in && g.interior(x, y, cz)));
        hr[(K + 1) & 3] = row_raw(x + 1);
        hc[(K + 1) & 3] = col_raw(x + 1);
        __syncthreads();
        if (!inyz) return;
        const int64_t px = (int64_t)x * P;
        const AdjPt<T>& m = pt[(K + 3) & 3];
        const AdjPt<T>& c = pt[K];
        const AdjPt<T>& p = pt[(K + 1) & 3];
        const int r = ty + 1, q = tz + 1;
        T oub[3] = {T(0), T(0), T(0)}, odb[3] = {T(0), T(0), T(0)};
        if (g.interior(x, y, z)) {
            double gmi = 0.0, gei = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T mw0 = c.mw[k];
                const T lmw = a.cx * (p.mw[k] + m.mw[k] - T(2) * mw0) +
                              a.cy * (L[k][r + 1][q] + L[k][r - 1][q] - T(2) * mw0) +
                              a.cz * (L[k][r][q + 1] + L[k][r][q - 1] - T(2) * mw0);
                oub[k] = c.ub[k] + lmw;
                const T u0 = c.u[k];
                const T lu = a.cx * (p.u[k] + m.u[k] - T(2) * u0) + a.cy * (L[3 + k][r + 1][q] + L[3 + k][r - 1][q] - T(2) * u0) +
                             a.cz * (L[3 + k][r][q + 1] + L[3 + k][r][q - 1] - T(2) * u0);
                gmi += (double)c.w[k] * (double)lu;
                T dbv = c.w[k];
                if (VISC) {
                    const T ew0 = c.ew[k];
                    const T lew = a.vx * (p.ew[k] + m.ew[k] - T(2) * ew0) +
                                  a.vy * (L[6 + k][r + 1][q] + L[6 + k][r - 1][q] - T(2) * ew0) +
                                  a.vz * (L[6 + k][r][q + 1] + L[6 + k][r][q - 1] - T(2) * ew0);
                    dbv += lew;
                    const T d0 = c.d[k];
                    const T ld = a.vx * (p.d[k] + m.d[k] - T(2) * d0) +
                                 a.vy * (L[9 + k][r + 1][q] + L[9 + k][r - 1][q] - T(2) * d0) +
                                 a.vz * (L[9 + k][r][q + 1] + L[9 + k][r][q - 1] - T(2) * d0);
                    gei += (double)c.w[k] * (double)ld;
                }
                odb[k] = dbv;
            }
            const uint32_t o8 = yz_el * 8u;
            auto& gm = at_byte(a.gm + px, o8);
            gm = gm + gmi;
            if (VISC) {
                auto& ge = at_byte(a.ge + px, o8);
                ge = ge + gei;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            at_byte(a.ub[k] + px, yz) = oub[k];
            at_byte(a.db[k] + px, yz) = odb[k];
        }
    });
}

// f[idx[k]] += grad[k]; an index outside [0, n) is skipped.
template <typename T>
__global__ void __launch_bounds__(256) inject_kernel(T* __restrict__ f, int64_t n, const int64_t* __restrict__ idx,
                                                     int64_t count, const T* __restrict__ grad, int64_t stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        const int64_t j = idx[k];
        if (j >= 0 && j < n) f[j] += grad[k * stride];
    }
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
        // the x-marching tile where it fits (one 64-wide z tile at least, planes of
        // fewer than 2^28 voxels for its 32-bit offsets), else one voxel per thread
        const char* env_k = getenv("SW_ADJ_KERNEL");
        constexpr int BY = adj_by<T>();
        bool march = nz >= 32 && ny >= BY && g.plane() < kMaxMarchPlane;
        if (env_k && env_k[0] == 'd') march = false;
        if (march) {
            auto launch = [&](auto kern) {
                static const int res = [&] {
                    int n = 0;
                    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, 64 * BY, 0) != hipSuccess || n < 1)
                        n = 1;
                    return n;
                }();
                const int nty = (int)((ny + BY - 1) / BY), ntz = (int)((nz + 63) / 64);
                const int64_t cols = (int64_t)nty * ntz, target = (int64_t)device_cu_count() * res;
                int64_t ntx = (target + cols - 1) / cols;
                if (ntx < 1) ntx = 1;
                int64_t xc = (nx + ntx - 1) / ntx;
                if (xc < 4) xc = 4;
                ntx = (nx + xc - 1) / xc;
                hipLaunchKernelGGL(kern, dim3((unsigned)(cols * ntx)), dim3(64, BY), 0, (hipStream_t)stream, a,
                                   (int)xc, nty, ntz);
            };
            if (eta_over_rho)
                launch(adjoint_march_kernel<T, true, BY>);
            else
                launch(adjoint_march_kernel<T, false, BY>);
        } else if (eta_over_rho) {
            hipLaunchKernelGGL((adjoint_step_kernel<T, true>), dim3(row_grid(g, 0)), row_block(), 0,
                               (hipStream_t)stream, a);
        } else {
            hipLaunchKernelGGL((adjoint_step_kernel<T, false>), dim3(row_grid(g, 0)), row_block(), 0,
                               (hipStream_t)stream, a);
        }
    });
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_inject_traces(int dtype, void* f, int64_t n, const int64_t* idx, int64_t count, const void* grad,
                     int64_t stride, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    SW_REQUIRE(count >= 0 && stride >= 1 && n >= 0, SW_ERR_ARG, "bad count/stride/n");
    if (count == 0) return SW_OK;
    SW_REQUIRE(f && idx && grad, SW_ERR_ARG, "null pointer");
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(inject_kernel<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                                   (hipStream_t)stream, (T*)f, n, idx, count, (const T*)grad, stride));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
