// Fused time step of the reduced Kelvin-Voigt shear model on gfx950.
//
// Restates one iteration of the loop in shear_fdtd (solver.py:194-234):
//   a      = mu/rho lap7(u) + eta/rho lap7(v) + f_S env[it] / rho,  v = mask (u - u_prev)/dT
//   u_next = u + (1 - vel_damp) (u - u_prev + dT^2 a),  masked to the interior
// as ONE kernel per step: both 7-point Laplacians (solver.py:686-695), the
// Kelvin-Voigt term, the force, damping, mask and update read each field
// once from HBM.
//
// State forms. SW_FORM_INCREMENT stores u_n and d_n = u_n - u_{n-1}:
//   d_{n+1} = mask (1 - vel_damp)(d_n + dT^2 mu/rho lap u_n + dT eta/rho lap d_n + env dT^2 f/rho)
//   u_{n+1} = u_n + d_{n+1}
// which equals the reference update in exact arithmetic, but in float32
// storage keeps the velocity d/dT to float32 precision instead of recovering
// it from two nearly equal rounded displacements. SW_FORM_LEAPFROG stores
// u_n and u_{n-1} like the reference (solver.py:148-150, 234).
//
// Two kernels:
//  * direct: one voxel per thread, all neighbours through the caches
//    (reference implementation on the device, small grids);
//  * march: a (BY x 64) tile of (y, z) columns marches along x (the slowest
//    axis) with a register queue of the x-1, x, x+1 planes; the y/z
//    neighbours of each plane come from a double-buffered LDS tile with a
//    one-voxel halo, so every field is read from HBM about once per step.
//    Tiles are assigned XCD-aware: consecutive tiles (which share halos) land
//    on the same XCD and therefore the same L2.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "sw_common.h"

namespace sw {
namespace {

template <typename T>
struct StepArgs {
    const T* A[3];  // u_n
    const T* B[3];  // d_n (increment) or u_{n-1} (leapfrog)
    T* Ao[3];       // u_{n+1}
    T* Bo[3];       // d_{n+1} (increment form)
    const T* m;     // mu / rho
    const T* e;     // eta / rho
    const T* f[3];  // dT^2 f_S / rho
    Grid g;
    T cx, cy, cz;  // dT^2 / h^2 per axis
    T vx, vy, vz;  // dT / h^2 per axis
    T damp;        // 1 - vel_damp
    T fs;          // env[it]
};

template <typename T, bool INCR, bool VISC, bool FORCE>
__global__ void __launch_bounds__(256) step_direct_kernel(StepArgs<T> a) {
    const Grid g = a.g;
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    SW_FOR_EACH_VOXEL(g, {
        T ou[3] = {T(0), T(0), T(0)}, od[3] = {T(0), T(0), T(0)};
        if (g.interior(x, y, z)) {
            const T mm = a.m[i];
            const T ee = VISC ? a.e[i] : T(0);
            for (int c = 0; c < 3; ++c) {
                const T* U = a.A[c];
                const T* B = a.B[c];
                auto dv = [&](int64_t k) { return INCR ? B[i + k] : (U[i + k] - B[i + k]); };
                const T u0 = U[i];
                const T lu = a.cx * (U[i + P] + U[i - P] - T(2) * u0) + a.cy * (U[i + R] + U[i - R] - T(2) * u0) +
                             a.cz * (U[i + 1] + U[i - 1] - T(2) * u0);
                const T d0 = dv(0);
                T inc = d0 + mm * lu;
                if (VISC) {
                    const T ld = a.vx * (dv(P) + dv(-P) - T(2) * d0) + a.vy * (dv(R) + dv(-R) - T(2) * d0) +
                                 a.vz * (dv(1) + dv(-1) - T(2) * d0);
                    inc += ee * ld;
                }
                if (FORCE) inc += a.fs * a.f[c][i];
                inc *= a.damp;
                ou[c] = u0 + inc;
                od[c] = inc;
            }
        }
        for (int c = 0; c < 3; ++c) {
            a.Ao[c][i] = ou[c];
            if (INCR) a.Bo[c][i] = od[c];
        }
    });
}

// One (y, z) column's state at one x plane: u and d = u - u_prev.
template <typename T>
struct Col {
    T u0, u1, u2, d0, d1, d2;
};

// Per-voxel coefficients: mu/rho, eta/rho and the force.
template <typename T>
struct Coef {
    T m, e, f0, f1, f2;
};

// Loads are unconditional from a clamped offset (voxel 0 when `ok` is false).
// Values loaded for points outside the grid only ever feed boundary voxels,
// whose outputs are overwritten with zero, so no select follows the load and
// the load's latency stays hidden behind the next barrier. (Voxel 0 is on the
// boundary layer, which holds zero in every state array.)
template <typename T, bool INCR>
__device__ __forceinline__ Col<T> load_col(const StepArgs<T>& a, int64_t o, bool ok) {
    const int64_t q = ok ? o : 0;
    const T u0 = a.A[0][q], u1 = a.A[1][q], u2 = a.A[2][q];
    const T b0 = a.B[0][q], b1 = a.B[1][q], b2 = a.B[2][q];
    Col<T> c;
    c.u0 = u0;
    c.u1 = u1;
    c.u2 = u2;
    c.d0 = INCR ? b0 : u0 - b0;
    c.d1 = INCR ? b1 : u1 - b1;
    c.d2 = INCR ? b2 : u2 - b2;
    return c;
}

template <typename T, bool VISC, bool FORCE>
__device__ __forceinline__ Coef<T> load_coef(const StepArgs<T>& a, int64_t o, bool ok) {
    const int64_t q = ok ? o : 0;
    Coef<T> k;
    k.m = a.m[q];
    k.e = VISC ? a.e[q] : T(0);
    k.f0 = FORCE ? a.f[0][q] : T(0);
    k.f1 = FORCE ? a.f[1][q] : T(0);
    k.f2 = FORCE ? a.f[2][q] : T(0);
    return k;
}

template <typename T, bool VISC>
__device__ __forceinline__ T comp_inc(const StepArgs<T>& a, T u0, T um, T up, T uyp, T uym, T uzp, T uzm, T d0, T dm,
                                      T dp, T dyp, T dym, T dzp, T dzm, T m, T e, T f) {
    const T lu = a.cx * (up + um - T(2) * u0) + a.cy * (uyp + uym - T(2) * u0) + a.cz * (uzp + uzm - T(2) * u0);
    T inc = d0 + m * lu + f;
    if (VISC) {
        const T ld = a.vx * (dp + dm - T(2) * d0) + a.vy * (dyp + dym - T(2) * d0) + a.vz * (dzp + dzm - T(2) * d0);
        inc += e * ld;
    }
    return inc * a.damp;
}

template <typename T, bool INCR, bool VISC, bool FORCE, int BY>
__global__ void __launch_bounds__(64 * BY) step_march_kernel(StepArgs<T> a, int xchunk, int nty, int ntz) {
    constexpr int BZ = 64;
    constexpr int NL = VISC ? 6 : 3;  // planes staged in LDS: u (3) and, with viscosity, d (3)
    constexpr int LY = BY + 2;
    constexpr int LZ = BZ + 2;
    __shared__ T lds[2][NL][LY][LZ];

    const Grid g = a.g;
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * BZ + tz;

    // XCD-aware tile order: the hardware deals blocks round-robin to the 8 XCDs.
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);
    const int tyi = (int)(b % (unsigned)nty);
    const unsigned rest = b / (unsigned)nty;
    const int tzi = (int)(rest % (unsigned)ntz);
    const int txi = (int)(rest / (unsigned)ntz);
    const int y0 = tyi * BY, z0 = tzi * BZ, x0 = txi * xchunk;
    const int x1 = min(x0 + xchunk, g.nx);
    const int y = y0 + ty, z = z0 + tz;
    const bool inyz = (y < g.ny) && (z < g.nz);
    const int64_t P = g.plane();
    const int64_t yz = inyz ? (int64_t)y * g.nz + z : 0;
    const bool yz_interior = inyz && y > 0 && y < g.ny - 1 && z > 0 && z < g.nz - 1;

    // halo points of the tile: two rows of 64 (y0-1, y0+BY) and two columns of BY (z0-1, z0+64)
    int hy = -1, hz = -1, ly = 0, lz = 0;
    if (tid < 64) {
        hy = y0 - 1, hz = z0 + tid, ly = 0, lz = tid + 1;
    } else if (tid < 128) {
        hy = y0 + BY, hz = z0 + tid - 64, ly = BY + 1, lz = tid - 63;
    } else if (tid < 128 + BY) {
        hy = y0 + tid - 128, hz = z0 - 1, ly = tid - 127, lz = 0;
    } else if (tid < 128 + 2 * BY) {
        hy = y0 + tid - 128 - BY, hz = z0 + BZ, ly = tid - 127 - BY, lz = BZ + 1;
    }
    const bool halo_thread = tid < 128 + 2 * BY;
    const bool halo_in = halo_thread && hy >= 0 && hy < g.ny && hz >= 0 && hz < g.nz;
    const int64_t hyz = halo_in ? (int64_t)hy * g.nz + hz : 0;

    auto center = [&](int xx) { return load_col<T, INCR>(a, (int64_t)xx * P + yz, inyz && xx >= 0 && xx < g.nx); };
    auto halo = [&](int xx) {
        Col<T> c{};
        if (halo_thread) c = load_col<T, INCR>(a, (int64_t)xx * P + hyz, halo_in && xx >= 0 && xx < g.nx);
        return c;
    };
    auto coef = [&](int xx) {
        return load_coef<T, VISC, FORCE>(a, (int64_t)xx * P + yz, yz_interior && xx > 0 && xx < g.nx - 1);
    };

    // One x plane. m, c, p: planes x-1, x, x+1 of this column; n receives x+2.
    // hc: this plane's halo point, hn receives the next one; kc / kn likewise for coefficients.
    auto plane = [&](auto buf, const Col<T>& m, const Col<T>& c, const Col<T>& p, Col<T>& n, const Col<T>& hc,
                     Col<T>& hn, const Coef<T>& kc, Coef<T>& kn, int x) {
        constexpr int B = decltype(buf)::value;
        auto& L = lds[B];
        L[0][ty + 1][tz + 1] = c.u0;
        L[1][ty + 1][tz + 1] = c.u1;
        L[2][ty + 1][tz + 1] = c.u2;
        if (VISC) {
            L[NL - 3][ty + 1][tz + 1] = c.d0;
            L[NL - 2][ty + 1][tz + 1] = c.d1;
            L[NL - 1][ty + 1][tz + 1] = c.d2;
        }
        if (halo_thread) {
            L[0][ly][lz] = hc.u0;
            L[1][ly][lz] = hc.u1;
            L[2][ly][lz] = hc.u2;
            if (VISC) {
                L[NL - 3][ly][lz] = hc.d0;
                L[NL - 2][ly][lz] = hc.d1;
                L[NL - 1][ly][lz] = hc.d2;
            }
        }
        n = center(x + 2);
        hn = halo(x + 1);
        kn = coef(x + 1);
        __syncthreads();
        if (!inyz) return;
        const bool in = yz_interior && x > 0 && x < g.nx - 1;
        const int r = ty + 1, q = tz + 1;
        auto nb = [&](int f, int dy, int dz) { return L[f][r + dy][q + dz]; };
        T i0 = comp_inc<T, VISC>(a, c.u0, m.u0, p.u0, nb(0, 1, 0), nb(0, -1, 0), nb(0, 0, 1), nb(0, 0, -1), c.d0, m.d0,
                                 p.d0, VISC ? nb(NL - 3, 1, 0) : T(0), VISC ? nb(NL - 3, -1, 0) : T(0),
                                 VISC ? nb(NL - 3, 0, 1) : T(0), VISC ? nb(NL - 3, 0, -1) : T(0), kc.m, kc.e,
                                 a.fs * kc.f0);
        T i1 = comp_inc<T, VISC>(a, c.u1, m.u1, p.u1, nb(1, 1, 0), nb(1, -1, 0), nb(1, 0, 1), nb(1, 0, -1), c.d1, m.d1,
                                 p.d1, VISC ? nb(NL - 2, 1, 0) : T(0), VISC ? nb(NL - 2, -1, 0) : T(0),
                                 VISC ? nb(NL - 2, 0, 1) : T(0), VISC ? nb(NL - 2, 0, -1) : T(0), kc.m, kc.e,
                                 a.fs * kc.f1);
        T i2 = comp_inc<T, VISC>(a, c.u2, m.u2, p.u2, nb(2, 1, 0), nb(2, -1, 0), nb(2, 0, 1), nb(2, 0, -1), c.d2, m.d2,
                                 p.d2, VISC ? nb(NL - 1, 1, 0) : T(0), VISC ? nb(NL - 1, -1, 0) : T(0),
                                 VISC ? nb(NL - 1, 0, 1) : T(0), VISC ? nb(NL - 1, 0, -1) : T(0), kc.m, kc.e,
                                 a.fs * kc.f2);
        if (!in) i0 = i1 = i2 = T(0);
        const int64_t o = (int64_t)x * P + yz;
        a.Ao[0][o] = in ? c.u0 + i0 : T(0);
        a.Ao[1][o] = in ? c.u1 + i1 : T(0);
        a.Ao[2][o] = in ? c.u2 + i2 : T(0);
        if (INCR) {
            a.Bo[0][o] = i0;
            a.Bo[1][o] = i1;
            a.Bo[2][o] = i2;
        }
    };

    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    Col<T> q0 = center(x0 - 1), q1 = center(x0), q2 = center(x0 + 1), q3;
    Col<T> h0 = halo(x0), h1;
    Coef<T> k0 = coef(x0), k1;
    // unrolled by four so the queue's roles rotate without register moves
    for (int x = x0; x < x1; x += 4) {
        plane(B0{}, q0, q1, q2, q3, h0, h1, k0, k1, x);
        if (x + 1 >= x1) break;
        plane(B1{}, q1, q2, q3, q0, h1, h0, k1, k0, x + 1);
        if (x + 2 >= x1) break;
        plane(B0{}, q2, q3, q0, q1, h0, h1, k0, k1, x + 2);
        if (x + 3 >= x1) break;
        plane(B1{}, q3, q0, q1, q2, h1, h0, k1, k0, x + 3);
    }
}

struct MarchShape {
    int by, xchunk, nty, ntz, ntx;
    unsigned blocks() const { return (unsigned)nty * ntz * ntx; }
};

MarchShape march_shape(const Grid& g, int dtype, int tile_y, int x_chunk) {
    MarchShape s;
    s.by = tile_y > 0 ? tile_y : (dtype == SW_F32 ? 8 : 4);
    s.nty = (g.ny + s.by - 1) / s.by;
    s.ntz = (g.nz + 63) / 64;
    if (x_chunk > 0) {
        s.xchunk = x_chunk;
    } else {
        // enough tiles for ~8 resident blocks per CU, chunks of at least 8 planes
        const int64_t target = (int64_t)device_cu_count() * 8;
        const int64_t cols = (int64_t)s.nty * s.ntz;
        int64_t ntx = (target + cols - 1) / cols;
        if (ntx < 1) ntx = 1;
        int64_t xc = (g.nx + ntx - 1) / ntx;
        if (xc < 8) xc = 8;
        s.xchunk = (int)xc;
    }
    s.ntx = (g.nx + s.xchunk - 1) / s.xchunk;
    return s;
}

template <typename T, bool INCR, bool VISC, bool FORCE>
int launch_step(const StepArgs<T>& a, int kernel, int tile_y, int x_chunk, hipStream_t s) {
    const Grid& g = a.g;
    if (kernel == SW_KERNEL_DIRECT) {
        hipLaunchKernelGGL((step_direct_kernel<T, INCR, VISC, FORCE>), dim3(row_grid(g, 0)), row_block(), 0, s, a);
        SW_LAUNCH_CHECK();
        return SW_OK;
    }
    const MarchShape ms = march_shape(g, DType<T>::code, tile_y, x_chunk);
    const dim3 grid(ms.blocks());
    switch (ms.by) {
        case 4:
            hipLaunchKernelGGL((step_march_kernel<T, INCR, VISC, FORCE, 4>), grid, dim3(64, 4), 0, s, a, ms.xchunk,
                               ms.nty, ms.ntz);
            break;
        case 8:
            hipLaunchKernelGGL((step_march_kernel<T, INCR, VISC, FORCE, 8>), grid, dim3(64, 8), 0, s, a, ms.xchunk,
                               ms.nty, ms.ntz);
            break;
        default:
            sw::set_error("tile_y %d not supported (4 or 8)", ms.by);
            return SW_ERR_ARG;
    }
    SW_LAUNCH_CHECK();
    return SW_OK;
}

template <typename T>
int launch_step_any(const StepArgs<T>& a, bool incr, bool visc, bool force, int kernel, int ty, int xc,
                    hipStream_t s) {
#define SW_STEP_CASE(I, V, F) \
    if (incr == I && visc == V && force == F) return launch_step<T, I, V, F>(a, kernel, ty, xc, s);
    SW_STEP_CASE(true, true, true)
    SW_STEP_CASE(true, true, false)
    SW_STEP_CASE(true, false, true)
    SW_STEP_CASE(true, false, false)
    SW_STEP_CASE(false, true, true)
    SW_STEP_CASE(false, true, false)
    SW_STEP_CASE(false, false, true)
    SW_STEP_CASE(false, false, false)
#undef SW_STEP_CASE
    return SW_ERR_ARG;
}

template <typename T>
int run_steps(sw_fdtd_state* st, int64_t n_steps, const double* env, const sw_fdtd_outputs* out, hipStream_t s) {
    const Grid g{(int)st->nx, (int)st->ny, (int)st->nz};
    const int64_t n = g.size();
    const bool incr = st->form == SW_FORM_INCREMENT;
    const bool visc = st->eta_over_rho != nullptr;
    const bool have_force = st->force[0] && st->force[1] && st->force[2];
    int kernel = st->kernel;
    if (kernel == SW_KERNEL_AUTO) kernel = (g.nz >= 32 && g.ny >= 8 && g.nx >= 8) ? SW_KERNEL_MARCH : SW_KERNEL_DIRECT;
    const int64_t step0 = out ? out->step0 : 0;

    StepArgs<T> a{};
    a.m = (const T*)st->mu_over_rho;
    a.e = (const T*)st->eta_over_rho;
    for (int c = 0; c < 3; ++c) a.f[c] = (const T*)st->force[c];
    a.g = g;
    const double dt = st->dt;
    a.cx = (T)(dt * dt / (st->dx * st->dx));
    a.cy = (T)(dt * dt / (st->dy * st->dy));
    a.cz = (T)(dt * dt / (st->dz * st->dz));
    a.vx = (T)(dt / (st->dx * st->dx));
    a.vy = (T)(dt / (st->dy * st->dy));
    a.vz = (T)(dt / (st->dz * st->dz));
    a.damp = (T)(1.0 - st->vel_damp);

    for (int64_t it = 0; it < n_steps; ++it) {
        const double fs = env ? env[it] : 0.0;
        const bool force = fs != 0.0;
        SW_REQUIRE(!force || have_force, SW_ERR_ARG, "env[%lld] != 0 but the state has no force", (long long)it);
        a.fs = (T)fs;
        for (int c = 0; c < 3; ++c) {
            a.A[c] = (const T*)st->u[c];
            a.B[c] = (const T*)st->b[c];
            a.Ao[c] = (T*)st->u_next[c];
            a.Bo[c] = (T*)st->b_next[c];
        }
        int rc = launch_step_any<T>(a, incr, visc, force, kernel, st->tile_y, st->x_chunk, s);
        if (rc) return rc;
        for (int c = 0; c < 3; ++c) {
            if (incr) {
                std::swap(st->u[c], st->u_next[c]);
                std::swap(st->b[c], st->b_next[c]);
            } else {  // (u_prev, u_curr) <- (u_curr, u_next); the old u_prev buffer is reused
                void* old_b = st->b[c];
                st->b[c] = st->u[c];
                st->u[c] = st->u_next[c];
                st->u_next[c] = old_b;
            }
        }
        if (!out) continue;
        const int64_t step = step0 + it;
        if (out->snap_stride > 0 && (step + 1) % out->snap_stride == 0) {
            const int64_t k = (step + 1) / out->snap_stride - 1;
            rc = sw_soa_to_aos3(DType<T>::code, st->u[0], st->u[1], st->u[2], (T*)out->snap_out + k * 3 * n, n, s);
            if (rc) return rc;
        }
        if (out->inv_stride > 0 && (step + 1) % out->inv_stride == 0) {
            const int64_t k = (step + 1) / out->inv_stride - 1;
            T* base = (T*)out->inv_out + k * 4 * n;
            rc = sw_strain_invariants_from_u(DType<T>::code, st->u[0], st->u[1], st->u[2], base, base + n,
                                             base + 2 * n, base + 3 * n, g.nx, g.ny, g.nz, st->dx, st->dy, st->dz,
                                             s);
            if (rc) return rc;
        }
        if (out->n_traces > 0) {
            rc = sw_gather_traces(DType<T>::code, st->u[2], out->trace_idx, out->n_traces,
                                  (T*)out->trace_out + step, out->trace_ld, s);
            if (rc) return rc;
        }
    }
    return SW_OK;
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" int sw_fdtd_run(sw_fdtd_state* st, int64_t n_steps, const double* env, const sw_fdtd_outputs* out,
                           void* stream) {
    SW_REQUIRE(st, SW_ERR_ARG, "null state");
    int rc = check_dtype(st->dtype);
    if (rc) return rc;
    if ((rc = check_grid(st->nx, st->ny, st->nz))) return rc;
    SW_REQUIRE(st->form == SW_FORM_INCREMENT || st->form == SW_FORM_LEAPFROG, SW_ERR_ARG, "unknown form %d",
               st->form);
    SW_REQUIRE(st->kernel >= SW_KERNEL_AUTO && st->kernel <= SW_KERNEL_MARCH, SW_ERR_ARG, "unknown kernel %d",
               st->kernel);
    SW_REQUIRE(n_steps >= 0, SW_ERR_ARG, "n_steps must be >= 0");
    SW_REQUIRE(st->dt > 0 && st->dx > 0 && st->dy > 0 && st->dz > 0, SW_ERR_ARG, "dt and spacing must be > 0");
    SW_REQUIRE(st->mu_over_rho, SW_ERR_ARG, "mu_over_rho is null");
    for (int c = 0; c < 3; ++c) {
        SW_REQUIRE(st->u[c] && st->b[c] && st->u_next[c], SW_ERR_ARG, "null state buffer");
        SW_REQUIRE(st->form != SW_FORM_INCREMENT || st->b_next[c], SW_ERR_ARG, "increment form needs b_next");
    }
    if (out) {
        SW_REQUIRE(out->step0 >= 0, SW_ERR_ARG, "step0 must be >= 0");
        SW_REQUIRE(out->snap_stride <= 0 || out->snap_out, SW_ERR_ARG, "snap_out is null");
        SW_REQUIRE(out->inv_stride <= 0 || out->inv_out, SW_ERR_ARG, "inv_out is null");
        SW_REQUIRE(out->n_traces <= 0 || (out->trace_idx && out->trace_out), SW_ERR_ARG, "null trace buffers");
        SW_REQUIRE(out->n_traces <= 0 || out->trace_ld >= out->step0 + n_steps, SW_ERR_ARG,
                   "trace_ld %lld < last step %lld", (long long)out->trace_ld, (long long)(out->step0 + n_steps));
    }
    hipStream_t s = (hipStream_t)stream;
    if (st->dtype == SW_F32) return run_steps<float>(st, n_steps, env, out, s);
    return run_steps<double>(st, n_steps, env, out, s);
}
