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

#include "march2_step.h"
#include "march3_step.h"
#include "march_step.h"
#include "sw_common.h"

namespace sw {
namespace {

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

template <typename T, bool INCR, bool VISC, bool FORCE, int BY, int BZ>
__global__ void __launch_bounds__(BZ * BY) step_march_kernel(StepArgs<T> a, int xchunk, int nty, int ntz) {
    __shared__ T lds[2][VISC ? 6 : 3][BY + 2][BZ + 2];
    // XCD-aware tile order: the hardware deals blocks round-robin to the 8 XCDs;
    // renumber so that consecutive tiles (which share halos) land on one XCD.
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);
    // z fastest, then y, then x chunk
    const int tzi = (int)(b % (unsigned)ntz);
    const unsigned rest = b / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int txi = (int)(rest / (unsigned)nty);
    const int x0 = txi * xchunk;
    march_tile<T, INCR, VISC, FORCE, BY, BZ>(a, x0, min(x0 + xchunk, a.g.nx), tyi * BY, tzi * BZ, lds);
}

template <typename T, bool VISC, bool FORCE, int BY, bool MID = false, bool STRAIN = false>
__global__ void __launch_bounds__(64 * (BY + 3)) step2_march_kernel(StepArgs<T> a, int xchunk, int nty, int ntz) {
    __shared__ Step2Lds<T, VISC, BY> lds;
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);  // XCD-aware, as step_march_kernel
    const int tzi = (int)(b % (unsigned)ntz);
    const unsigned rest = b / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int txi = (int)(rest / (unsigned)nty);
    const int x0 = txi * xchunk;
    march2_tile<T, VISC, FORCE, BY, MID, 0, STRAIN>(a, x0, min(x0 + xchunk, a.g.nx), tyi * BY, tzi * 64, lds);
}

// SW_STEP3_BY: rows of the three-step tile (block of 64 x (BY + 5) threads).
#ifndef SW_STEP3_BY
#define SW_STEP3_BY 7
#endif

template <typename T, bool VISC, int BY>
__global__ void __launch_bounds__(64 * (BY + 5)) step3_march_kernel(StepArgs<T> a, int xchunk, int nty, int ntz) {
    __shared__ Step3Lds<T, VISC, BY> lds;
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);  // XCD-aware, as step_march_kernel
    const int tzi = (int)(b % (unsigned)ntz);
    const unsigned rest = b / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int txi = (int)(rest / (unsigned)nty);
    const int x0 = txi * xchunk;
    march3_tile<T, VISC, BY>(a, x0, min(x0 + xchunk, a.g.nx), tyi * BY, tzi * 64, lds);
}

struct MarchShape {
    int by, bz, xchunk, nty, ntz, ntx;
    unsigned blocks() const { return (unsigned)nty * ntz * ntx; }
};

// Launch shape: a (bz x by) tile of (y, z) columns and an x chunk that makes
// the grid one full wave of resident blocks (occupancy of this instantiation
// x CUs), unless the caller fixes it.
MarchShape march_shape(const Grid& g, int by, int bz, int x_chunk, int resident_per_cu) {
    MarchShape s;
    s.by = by;
    s.bz = bz;
    s.nty = (g.ny + by - 1) / by;
    s.ntz = (g.nz + bz - 1) / bz;
    if (x_chunk > 0) {
        s.xchunk = x_chunk;
    } else {
        const int64_t target = (int64_t)device_cu_count() * (resident_per_cu > 0 ? resident_per_cu : 2);
        const int64_t cols = (int64_t)s.nty * s.ntz;
        int64_t ntx = (target + cols / 2) / cols;
        if (ntx < 1) ntx = 1;
        int64_t xc = (g.nx + ntx - 1) / ntx;
        if (xc < 8) xc = 8;
        s.xchunk = (int)xc;
    }
    s.ntx = (g.nx + s.xchunk - 1) / s.xchunk;
    return s;
}

template <typename K>
int resident_blocks(K kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
    return n;
}

template <typename T, bool INCR, bool VISC, bool FORCE, int BY, int BZ>
void launch_march(const StepArgs<T>& a, int x_chunk, hipStream_t s) {
    auto k = step_march_kernel<T, INCR, VISC, FORCE, BY, BZ>;
    static const int res = resident_blocks(k, BY * BZ);
    const MarchShape ms = march_shape(a.g, BY, BZ, x_chunk, res);
    hipLaunchKernelGGL(k, dim3(ms.blocks()), dim3(BZ, BY), 0, s, a, ms.xchunk, ms.nty, ms.ntz);
}

template <typename T, bool INCR, bool VISC, bool FORCE>
int launch_step(const StepArgs<T>& a, int kernel, int tile_y, int tile_z, int x_chunk, hipStream_t s) {
    const Grid& g = a.g;
    if (kernel == SW_KERNEL_DIRECT) {
        hipLaunchKernelGGL((step_direct_kernel<T, INCR, VISC, FORCE>), dim3(row_grid(g, 0)), row_block(), 0, s, a);
        SW_LAUNCH_CHECK();
        return SW_OK;
    }
    // default tile: 8 x 64 (fp32) or 8 x 64 (fp64); full-width 256 rows on request
    const int bz = tile_z > 0 ? tile_z : 64;
    const int by = tile_y > 0 ? tile_y : (bz == 256 ? (DType<T>::code == SW_F32 ? 4 : 2) : 8);
    if (bz == 64 && by == 4) {
        launch_march<T, INCR, VISC, FORCE, 4, 64>(a, x_chunk, s);
    } else if (bz == 64 && by == 8) {
        launch_march<T, INCR, VISC, FORCE, 8, 64>(a, x_chunk, s);
    } else if (bz == 256 && by == 2) {
        launch_march<T, INCR, VISC, FORCE, 2, 256>(a, x_chunk, s);
    } else if (bz == 256 && by == 4 && DType<T>::code == SW_F32) {
        if constexpr (DType<T>::code == SW_F32) launch_march<T, INCR, VISC, FORCE, 4, 256>(a, x_chunk, s);
    } else {
        sw::set_error("tile %d x %d (y x z) not supported: 4x64, 8x64, 2x256, 4x256 (float32)", by, bz);
        return SW_ERR_ARG;
    }
    SW_LAUNCH_CHECK();
    return SW_OK;
}

template <typename T, bool VISC, bool FORCE, int BY, bool MID = false, bool STRAIN = false>
void launch_march2(const StepArgs<T>& a, int x_chunk, hipStream_t s) {
    auto k = step2_march_kernel<T, VISC, FORCE, BY, MID, STRAIN>;
    static const int res = resident_blocks(k, 64 * (BY + 3));
    const MarchShape ms = march_shape(a.g, BY, 64, x_chunk, res);
    hipLaunchKernelGGL(k, dim3(ms.blocks()), dim3(64, BY + 3), 0, s, a, ms.xchunk, ms.nty, ms.ntz);
}

// Two increment-form steps in one pass (u_n, d_n -> u_{n+2}, d_{n+2}); the
// force, when either step applies it, is scaled by a.fs and a.fs2. Tile rows:
// float32 8 (default) or 12 (force-free passes; forced ones use 8); float64 4
// (a 7-wave block leaves each wave the registers the float64 queues need).
// STRAIN (force-free passes only) also stores the strain invariants of u_n.
template <typename T, bool VISC, bool FORCE, bool MID = false, bool STRAIN = false>
int launch_step2_rows(const StepArgs<T>& a, int tile_y, int x_chunk, hipStream_t s) {
    static_assert(!STRAIN || (!FORCE && !MID), "the strain pass is force-free");
    if constexpr (DType<T>::code == SW_F32) {
        const int by = tile_y > 0 ? tile_y : 8;
        if (by == 8 || (by == 12 && (FORCE || MID))) {
            launch_march2<T, VISC, FORCE, 8, MID, STRAIN>(a, x_chunk, s);
        } else if (by == 12) {
            if constexpr (!FORCE && !MID) launch_march2<T, VISC, false, 12, false, STRAIN>(a, x_chunk, s);
        } else {
            sw::set_error("two-step pass: float32 tile_y %d not supported (8 or 12)", by);
            return SW_ERR_ARG;
        }
    } else {
        const int by = tile_y > 0 ? tile_y : 4;
        if (by != 4) {
            sw::set_error("two-step pass: float64 tile_y %d not supported (4)", by);
            return SW_ERR_ARG;
        }
        launch_march2<T, VISC, FORCE, 4, MID, STRAIN>(a, x_chunk, s);
    }
    SW_LAUNCH_CHECK();
    return SW_OK;
}

// Three force-free increment-form steps in one pass (u_n, d_n -> u_{n+3},
// d_{n+3}); float32 only (the float64 queues would not fit the registers of
// a (BY + 5)-wave block).
template <typename T>
int launch_step3(const StepArgs<T>& a, bool visc, int x_chunk, hipStream_t s) {
    if constexpr (DType<T>::code == SW_F32) {
        constexpr int BY = SW_STEP3_BY;
        auto launch = [&](auto kern) {
            static const int res = resident_blocks(kern, 64 * (BY + 5));
            const MarchShape ms = march_shape(a.g, BY, 64, x_chunk, res);
            hipLaunchKernelGGL(kern, dim3(ms.blocks()), dim3(64, BY + 5), 0, s, a, ms.xchunk, ms.nty, ms.ntz);
        };
        if (visc)
            launch(step3_march_kernel<T, true, BY>);
        else
            launch(step3_march_kernel<T, false, BY>);
        SW_LAUNCH_CHECK();
        return SW_OK;
    } else {
        sw::set_error("three-step pass: float32 only");
        return SW_ERR_UNSUPPORTED;
    }
}

template <typename T, bool MID = false>
int launch_step2(const StepArgs<T>& a, bool visc, bool force, int tile_y, int x_chunk, hipStream_t s,
                 bool strain = false) {
    if constexpr (!MID) {
        if (strain && !force)
            return visc ? launch_step2_rows<T, true, false, false, true>(a, tile_y, x_chunk, s)
                        : launch_step2_rows<T, false, false, false, true>(a, tile_y, x_chunk, s);
    }
    if (visc) return force ? launch_step2_rows<T, true, true, MID>(a, tile_y, x_chunk, s)
                           : launch_step2_rows<T, true, false, MID>(a, tile_y, x_chunk, s);
    return force ? launch_step2_rows<T, false, true, MID>(a, tile_y, x_chunk, s)
                 : launch_step2_rows<T, false, false, MID>(a, tile_y, x_chunk, s);
}

template <typename T>
int launch_step_any(const StepArgs<T>& a, bool incr, bool visc, bool force, int kernel, int ty, int tz, int xc,
                    hipStream_t s) {
#define SW_STEP_CASE(I, V, F) \
    if (incr == I && visc == V && force == F) return launch_step<T, I, V, F>(a, kernel, ty, tz, xc, s);
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

// Traces of a two-step pass (increment form), whose state is u_{n+2}, d_{n+2}:
// column c+1 gets u_z after step n+2, column c gets u_z - d_z, u_z after step
// n+1 (u_n + d_{n+1} up to one rounding of u). Indices outside [0, n) give NaN.
template <typename T>
__global__ void __launch_bounds__(256) gather2_kernel(const T* __restrict__ uz, const T* __restrict__ dz, int64_t n,
                                                      const int64_t* __restrict__ idx, int64_t count,
                                                      T* __restrict__ out, int64_t stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int64_t j = idx[k];
    const bool ok = j >= 0 && j < n;
    const T u = ok ? uz[j] : (T)NAN;
    out[k * stride] = ok ? u - dz[j] : (T)NAN;
    out[k * stride + 1] = u;
}

// Non-finite detection: if any of the three fields holds an inf or NaN and
// *flag is still -1, *flag = step. One compare-and-swap per wave that saw one.
template <typename T>
__global__ void __launch_bounds__(256) finite_check_kernel(const T* __restrict__ a, const T* __restrict__ b,
                                                           const T* __restrict__ c, int64_t n, int64_t step,
                                                           long long* flag) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        bad |= !(isfinite(a[i]) && isfinite(b[i]) && isfinite(c[i]));
    if (__any(bad) && (threadIdx.x & (kWave - 1)) == 0) atomicCAS((unsigned long long*)flag, ~0ull, (unsigned long long)step);
}

// The coefficients, force and step factors of a state (the buffers are set per pass).
template <typename T>
StepArgs<T> step_args(const sw_fdtd_state* st) {
    StepArgs<T> a{};
    a.m = (const T*)st->mu_over_rho;
    a.e = (const T*)st->eta_over_rho;
    for (int c = 0; c < 3; ++c) a.f[c] = (const T*)st->force[c];
    a.g = Grid{(int)st->nx, (int)st->ny, (int)st->nz};
    const double dt = st->dt;
    a.cx = (T)(dt * dt / (st->dx * st->dx));
    a.cy = (T)(dt * dt / (st->dy * st->dy));
    a.cz = (T)(dt * dt / (st->dz * st->dz));
    a.vx = (T)(dt / (st->dx * st->dx));
    a.vy = (T)(dt / (st->dy * st->dy));
    a.vz = (T)(dt / (st->dz * st->dz));
    a.damp = (T)(1.0 - st->vel_damp);
    return a;
}

template <typename T>
int run_steps(sw_fdtd_state* st, int64_t n_steps, const double* env, const sw_fdtd_outputs* out, hipStream_t s) {
    const Grid g{(int)st->nx, (int)st->ny, (int)st->nz};
    const int64_t n = g.size();
    const bool incr = st->form == SW_FORM_INCREMENT;
    const bool visc = st->eta_over_rho != nullptr;
    const bool have_force = st->force[0] && st->force[1] && st->force[2];
    int kernel = st->kernel;
    const bool march_fits = g.nz >= 32 && g.ny >= 8 && g.nx >= 8 && g.plane() < kMaxMarchPlane;
    if (kernel == SW_KERNEL_AUTO) kernel = march_fits ? SW_KERNEL_MARCH2 : SW_KERNEL_DIRECT;
    // three-step passes: float32 only; other storage runs two-step passes
    const bool triples = kernel == SW_KERNEL_MARCH3 && incr && DType<T>::code == SW_F32;
    if (kernel == SW_KERNEL_MARCH3) kernel = SW_KERNEL_MARCH2;
    SW_REQUIRE(kernel == SW_KERNEL_DIRECT || g.plane() < kMaxMarchPlane, SW_ERR_ARG,
               "march kernels need a y-z plane of fewer than %lld voxels", (long long)kMaxMarchPlane);
    // two-step passes (increment form, no output in between); the remaining
    // single steps of such a run use the one-step march kernel
    const bool pairs = kernel == SW_KERNEL_MARCH2 && incr;
    const int one_kernel = kernel == SW_KERNEL_MARCH2 ? SW_KERNEL_MARCH : kernel;
    // tile_y names the two-step tile in a two-step run; its single steps use the default tile
    const int one_tile_y = kernel == SW_KERNEL_MARCH2 ? 0 : st->tile_y;
    const int64_t step0 = out ? out->step0 : 0;
    // whether an output reads the whole state after global step `step` (a pass ends there); traces
    // do not end passes: a two-step pass gathers both of its steps' traces afterwards
    auto output_after = [&](int64_t step) {
        if (!out) return false;
        if (out->snap_stride > 0 && (step + 1) % out->snap_stride == 0) return true;
        return out->inv_stride > 0 && (step + 1) % out->inv_stride == 0;
    };
    const bool traces = out && out->n_traces > 0;

    StepArgs<T> a = step_args<T>(st);
    a.ihx = (T)(0.5 / st->dx);
    a.ihy = (T)(0.5 / st->dy);
    a.ihz = (T)(0.5 / st->dz);
    // In-loop strain invariants of the state after an output step are computed by the next pass when
    // that is a force-free two-step pass (it reads the state anyway); otherwise, and after the last
    // pass of this call, by their own kernel.
    const bool fuse_strain = pairs && march_fits && out && out->inv_mode == 0;
    T* inv_pending = nullptr;
    auto flush_inv = [&]() -> int {
        if (!inv_pending) return SW_OK;
        T* base = inv_pending;
        inv_pending = nullptr;
        return strain_invariants_in_loop(DType<T>::code, st->u[0], st->u[1], st->u[2], base, g.nx, g.ny, g.nz,
                                         st->dx, st->dy, st->dz, s);
    };

    for (int64_t it = 0; it < n_steps;) {
        const double fs = env ? env[it] : 0.0;
        // three force-free steps with no output after the first two
        const bool three = triples && !traces && it + 2 < n_steps && !output_after(step0 + it) &&
                           !output_after(step0 + it + 1) &&
                           (!env || (env[it] == 0.0 && env[it + 1] == 0.0 && env[it + 2] == 0.0));
        const bool two = !three && pairs && it + 1 < n_steps && !output_after(step0 + it);
        const double fs2 = (two && env) ? env[it + 1] : 0.0;
        const bool force = fs != 0.0 || fs2 != 0.0;
        SW_REQUIRE(!force || have_force, SW_ERR_ARG, "env[%lld] != 0 but the state has no force",
                   (long long)(fs != 0.0 ? it : it + 1));
        a.fs = (T)fs;
        a.fs2 = (T)fs2;
        for (int c = 0; c < 3; ++c) {
            a.A[c] = (const T*)st->u[c];
            a.B[c] = (const T*)st->b[c];
            a.Ao[c] = (T*)st->u_next[c];
            a.Bo[c] = (T*)st->b_next[c];
        }
        const bool strain = inv_pending && two && !force;
        int rc = strain ? SW_OK : flush_inv();
        if (rc) return rc;
        if (strain)
            for (int k = 0; k < 4; ++k) a.inv[k] = inv_pending + k * n;
        rc = three ? launch_step3<T>(a, visc, st->x_chunk, s)
             : two ? launch_step2<T>(a, visc, force, st->tile_y, st->x_chunk, s, strain)
                   : launch_step_any<T>(a, incr, visc, force, one_kernel, one_tile_y, st->tile_z, st->x_chunk, s);
        if (rc) return rc;
        if (strain) inv_pending = nullptr;
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
        const int64_t done_before = step0 + it;
        it += three ? 3 : two ? 2 : 1;
        if (!out) continue;
        const int64_t step = step0 + it - 1;  // the last step of this pass
        if (out->finite_stride > 0 && (step + 1) / out->finite_stride > done_before / out->finite_stride) {
            hipLaunchKernelGGL(finite_check_kernel<T>, dim3(flat_grid(n)), dim3(256), 0, s, (const T*)st->u[0],
                               (const T*)st->u[1], (const T*)st->u[2], n, step, (long long*)out->nonfinite_step);
            SW_LAUNCH_CHECK();
        }
        if (out->snap_stride > 0 && (step + 1) % out->snap_stride == 0) {
            const int64_t k = (step + 1) / out->snap_stride - 1;
            rc = sw_soa_to_aos3(DType<T>::code, st->u[0], st->u[1], st->u[2], (T*)out->snap_out + k * 3 * n, n, s);
            if (rc) return rc;
        }
        if (out->inv_stride > 0 && (step + 1) % out->inv_stride == 0) {
            const int64_t k = (step + 1) / out->inv_stride - 1;
            inv_pending = (T*)out->inv_out + k * 4 * n;
            if (!fuse_strain && (rc = flush_inv())) return rc;
        }
        if (traces && two) {
            hipLaunchKernelGGL(gather2_kernel<T>, dim3((unsigned)((out->n_traces + 255) / 256)), dim3(256), 0, s,
                               (const T*)st->u[2], (const T*)st->b[2], n, out->trace_idx, out->n_traces,
                               (T*)out->trace_out + step - 1, out->trace_ld);
            SW_LAUNCH_CHECK();
        } else if (traces) {
            rc = sw_gather_traces(DType<T>::code, st->u[2], n, out->trace_idx, out->n_traces,
                                  (T*)out->trace_out + step, out->trace_ld, s);
            if (rc) return rc;
        }
    }
    return flush_inv();
}

// One two-step pass from (u, b) into (u_next, b_next) that also stores its
// intermediate state into (u_mid, d_mid); the state's buffers are not rotated.
template <typename T>
int run_step2_mid(const sw_fdtd_state* st, const double* env2, void* const* u_mid, void* const* d_mid,
                  hipStream_t s) {
    StepArgs<T> a = step_args<T>(st);
    const double fs = env2 ? env2[0] : 0.0, fs2 = env2 ? env2[1] : 0.0;
    const bool force = fs != 0.0 || fs2 != 0.0;
    SW_REQUIRE(!force || (st->force[0] && st->force[1] && st->force[2]), SW_ERR_ARG,
               "env != 0 but the state has no force");
    a.fs = (T)fs;
    a.fs2 = (T)fs2;
    for (int c = 0; c < 3; ++c) {
        a.A[c] = (const T*)st->u[c];
        a.B[c] = (const T*)st->b[c];
        a.Ao[c] = (T*)st->u_next[c];
        a.Bo[c] = (T*)st->b_next[c];
        a.Am[c] = (T*)u_mid[c];
        a.Bm[c] = (T*)d_mid[c];
    }
    return launch_step2<T, true>(a, st->eta_over_rho != nullptr, force, st->tile_y, st->x_chunk, s);
}

// n_passes force-free two-step passes, each from (u, b) into (u_next,
// b_next) with the state's tile_y and x_chunk; the buffers are not rotated.
template <typename T>
int run_trial_passes(const sw_fdtd_state* st, int64_t n_passes, hipStream_t s) {
    StepArgs<T> a = step_args<T>(st);
    a.fs = a.fs2 = T(0);
    for (int c = 0; c < 3; ++c) {
        a.A[c] = (const T*)st->u[c];
        a.B[c] = (const T*)st->b[c];
        a.Ao[c] = (T*)st->u_next[c];
        a.Bo[c] = (T*)st->b_next[c];
    }
    const bool visc = st->eta_over_rho != nullptr;
    for (int64_t k = 0; k < n_passes; ++k)
        if (const int rc = launch_step2<T>(a, visc, false, st->tile_y, st->x_chunk, s)) return rc;
    return SW_OK;
}

// The checks shared by the entry points that run two-step passes on a state
// without rotating it.
int check_pass_state(const sw_fdtd_state* st) {
    SW_REQUIRE(st, SW_ERR_ARG, "null state");
    int rc = check_dtype(st->dtype);
    if (rc) return rc;
    if ((rc = check_grid(st->nx, st->ny, st->nz))) return rc;
    SW_REQUIRE(st->form == SW_FORM_INCREMENT, SW_ERR_ARG, "the two-step pass needs the increment form");
    SW_REQUIRE(st->dt > 0 && st->dx > 0 && st->dy > 0 && st->dz > 0, SW_ERR_ARG, "dt and spacing must be > 0");
    SW_REQUIRE(st->mu_over_rho, SW_ERR_ARG, "mu_over_rho is null");
    for (int c = 0; c < 3; ++c)
        SW_REQUIRE(st->u[c] && st->b[c] && st->u_next[c] && st->b_next[c], SW_ERR_ARG, "null state buffer");
    const Grid g{(int)st->nx, (int)st->ny, (int)st->nz};
    SW_REQUIRE(g.nz >= 32 && g.ny >= 8 && g.nx >= 8 && g.plane() < kMaxMarchPlane, SW_ERR_UNSUPPORTED,
               "the two-step pass needs nz >= 32, ny >= 8, nx >= 8 and a y-z plane of fewer than %lld voxels",
               (long long)kMaxMarchPlane);
    return SW_OK;
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" int sw_fdtd_trial_passes(const sw_fdtd_state* st, int64_t n_passes, void* stream) {
    if (const int rc = check_pass_state(st)) return rc;
    SW_REQUIRE(n_passes >= 0, SW_ERR_ARG, "n_passes must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    if (st->dtype == SW_F32) return run_trial_passes<float>(st, n_passes, s);
    return run_trial_passes<double>(st, n_passes, s);
}

extern "C" int sw_fdtd_step2_mid(const sw_fdtd_state* st, const double* env2, void* const* u_mid,
                                 void* const* d_mid, void* stream) {
    if (const int rc = check_pass_state(st)) return rc;
    SW_REQUIRE(u_mid && d_mid, SW_ERR_ARG, "null intermediate state");
    for (int c = 0; c < 3; ++c) SW_REQUIRE(u_mid[c] && d_mid[c], SW_ERR_ARG, "null intermediate state buffer");
    hipStream_t s = (hipStream_t)stream;
    if (st->dtype == SW_F32) return run_step2_mid<float>(st, env2, u_mid, d_mid, s);
    return run_step2_mid<double>(st, env2, u_mid, d_mid, s);
}

extern "C" int sw_fdtd_run(sw_fdtd_state* st, int64_t n_steps, const double* env, const sw_fdtd_outputs* out,
                           void* stream) {
    SW_REQUIRE(st, SW_ERR_ARG, "null state");
    int rc = check_dtype(st->dtype);
    if (rc) return rc;
    if ((rc = check_grid(st->nx, st->ny, st->nz))) return rc;
    SW_REQUIRE(st->form == SW_FORM_INCREMENT || st->form == SW_FORM_LEAPFROG, SW_ERR_ARG, "unknown form %d",
               st->form);
    SW_REQUIRE(st->kernel >= SW_KERNEL_AUTO && st->kernel <= SW_KERNEL_MARCH3, SW_ERR_ARG, "unknown kernel %d",
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
        SW_REQUIRE(out->finite_stride <= 0 || out->nonfinite_step, SW_ERR_ARG, "nonfinite_step is null");
        SW_REQUIRE(out->n_traces <= 0 || out->trace_ld >= out->step0 + n_steps, SW_ERR_ARG,
                   "trace_ld %lld < last step %lld", (long long)out->trace_ld, (long long)(out->step0 + n_steps));
    }
    hipStream_t s = (hipStream_t)stream;
    if (st->dtype == SW_F32) return run_steps<float>(st, n_steps, env, out, s);
    return run_steps<double>(st, n_steps, env, out, s);
}
