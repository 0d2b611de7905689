// Shared device code of the fused time step: the per-step arguments and one
// x-marching tile. Used by the single-solve kernel (fdtd.hip) and the batched
// persistent kernel (batch.hip), so a solve's arithmetic is identical in both.
#pragma once

#include <type_traits>

#include "sw_common.h"

namespace sw {

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

// Single-use streams (coefficients, force) are read with the non-temporal
// hint, so they do not evict halo lines from the L2 (SW_NO_NT: plain loads).
template <typename T>
__device__ __forceinline__ T load_once(const T* p) {
#ifndef SW_NO_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

template <typename T, bool VISC, bool FORCE>
__device__ __forceinline__ Coef<T> load_coef(const StepArgs<T>& a, int64_t o, bool ok) {
    const int64_t q = ok ? o : 0;
    Coef<T> k;
    k.m = load_once(a.m + q);
    k.e = VISC ? load_once(a.e + q) : T(0);
    k.f0 = FORCE ? load_once(a.f[0] + q) : T(0);
    k.f1 = FORCE ? load_once(a.f[1] + q) : T(0);
    k.f2 = FORCE ? load_once(a.f[2] + q) : T(0);
    return k;
}

// Output store of the time step: the next step's state is written with the
// non-temporal hint (it is re-read one launch later, after the L2 has turned
// over), leaving the L2 to halo reuse. SW_NO_NT builds plain stores.
template <typename T>
__device__ __forceinline__ void store_out(T* p, T v) {
#ifndef SW_NO_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
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

// One tile of the time step: columns (y0 .. y0+BY-1, z0 .. z0+BZ-1), planes
// x0 .. x1-1; the block is (BZ, BY) threads (BZ a multiple of 64, so a wave
// never spans two rows) and all of them call this together. `lds` holds at
// least NL = (VISC ? 6 : 3) planes of (BY+2) x (BZ+2).
template <typename T, bool INCR, bool VISC, bool FORCE, int BY, int BZ, int NLD>
__device__ __forceinline__ void march_tile(const StepArgs<T>& a, int x0, int x1, int y0, int z0,
                                           T (&lds)[2][NLD][BY + 2][BZ + 2]) {
    constexpr int NL = VISC ? 6 : 3;  // planes staged in LDS: u (3) and, with viscosity, d (3)
    static_assert(NLD >= NL, "LDS too small");
    static_assert(BZ % 64 == 0 && BY >= 2, "tile shape");
    const Grid g = a.g;
    const int tz = threadIdx.x, ty = threadIdx.y;

    const int y = y0 + ty, z = z0 + tz;
    const bool inyz = (y < g.ny) && (z < g.nz);
    const int64_t P = g.plane();
    const int64_t yz = inyz ? (int64_t)y * g.nz + z : 0;
    const bool yz_interior = inyz && y > 0 && y < g.ny - 1 && z > 0 && z < g.nz - 1;

    // Halo of the tile. Row halos (y0-1, y0+BY) are loaded by the first and
    // last row of threads, at their own z (wave-uniform); column halos
    // (z0-1, z0+BZ) by the first and last lane of each row. A halo point
    // outside the grid is read from voxel 0 (on the boundary layer, zero) and
    // only feeds boundary voxels.
    const bool row_thread = ty == 0 || ty == BY - 1;
    const int ry = ty == 0 ? y0 - 1 : y0 + BY;
    const int rly = ty == 0 ? 0 : BY + 1;
    const bool row_in = row_thread && ry >= 0 && ry < g.ny && z < g.nz;
    const int64_t ryz = row_in ? (int64_t)ry * g.nz + z : 0;
    const bool col_thread = tz == 0 || tz == BZ - 1;
    const int cz = tz == 0 ? z0 - 1 : z0 + BZ;
    const int clz = tz == 0 ? 0 : BZ + 1;
    const bool col_in = col_thread && cz >= 0 && cz < g.nz && y < g.ny;
    const int64_t cyz = col_in ? (int64_t)y * g.nz + cz : 0;

    auto center = [&](int xx) { return load_col<T, INCR>(a, (int64_t)xx * P + yz, inyz && xx >= 0 && xx < g.nx); };
    auto row_halo = [&](int xx) {
        Col<T> c{};
        if (row_thread) c = load_col<T, INCR>(a, (int64_t)xx * P + ryz, row_in && xx >= 0 && xx < g.nx);
        return c;
    };
    auto col_halo = [&](int xx) {
        Col<T> c{};
        if (col_thread) c = load_col<T, INCR>(a, (int64_t)xx * P + cyz, col_in && xx >= 0 && xx < g.nx);
        return c;
    };
    auto coef = [&](int xx) {
        return load_coef<T, VISC, FORCE>(a, (int64_t)xx * P + yz, yz_interior && xx > 0 && xx < g.nx - 1);
    };

    // One x plane. m, c, p: planes x-1, x, x+1 of this column; n receives x+2.
    // hc: this plane's halo point, hn receives the next one; kc / kn likewise for coefficients.
    auto plane = [&](auto buf, const Col<T>& m, const Col<T>& c, const Col<T>& p, Col<T>& n, const Col<T>& hc,
                     Col<T>& hn, const Col<T>& vc, Col<T>& vn, const Coef<T>& kc, Coef<T>& kn, int x) {
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
        if (row_thread) {
            L[0][rly][tz + 1] = hc.u0;
            L[1][rly][tz + 1] = hc.u1;
            L[2][rly][tz + 1] = hc.u2;
            if (VISC) {
                L[NL - 3][rly][tz + 1] = hc.d0;
                L[NL - 2][rly][tz + 1] = hc.d1;
                L[NL - 1][rly][tz + 1] = hc.d2;
            }
        }
        if (col_thread) {
            L[0][ty + 1][clz] = vc.u0;
            L[1][ty + 1][clz] = vc.u1;
            L[2][ty + 1][clz] = vc.u2;
            if (VISC) {
                L[NL - 3][ty + 1][clz] = vc.d0;
                L[NL - 2][ty + 1][clz] = vc.d1;
                L[NL - 1][ty + 1][clz] = vc.d2;
            }
        }
        n = center(x + 2);
        hn = row_halo(x + 1);
        vn = col_halo(x + 1);
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
        store_out(a.Ao[0] + o, in ? c.u0 + i0 : T(0));
        store_out(a.Ao[1] + o, in ? c.u1 + i1 : T(0));
        store_out(a.Ao[2] + o, in ? c.u2 + i2 : T(0));
        if (INCR) {
            store_out(a.Bo[0] + o, i0);
            store_out(a.Bo[1] + o, i1);
            store_out(a.Bo[2] + o, i2);
        }
    };

    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    Col<T> q0 = center(x0 - 1), q1 = center(x0), q2 = center(x0 + 1), q3;
    Col<T> h0 = row_halo(x0), h1, c0 = col_halo(x0), c1;
    Coef<T> k0 = coef(x0), k1;
    // unrolled by four so the queue's roles rotate without register moves
    for (int x = x0; x < x1; x += 4) {
        plane(B0{}, q0, q1, q2, q3, h0, h1, c0, c1, k0, k1, x);
        if (x + 1 >= x1) break;
        plane(B1{}, q1, q2, q3, q0, h1, h0, c1, c0, k1, k0, x + 1);
        if (x + 2 >= x1) break;
        plane(B0{}, q2, q3, q0, q1, h0, h1, c0, c1, k0, k1, x + 2);
        if (x + 3 >= x1) break;
        plane(B1{}, q3, q0, q1, q2, h1, h0, c1, c0, k1, k0, x + 3);
    }
}

}  // namespace sw
