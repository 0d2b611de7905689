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
    T fs2;         // env[it+1] (the second step of a two-step pass)
    T* Am[3];      // u_{n+1} of a two-step pass (its intermediate state, when asked for)
    T* Bm[3];      // d_{n+1} of a two-step pass
    T* inv[4];     // the strain invariants of u_n (volumetric, deviatoric, max shear, von Mises), when asked for
    T ihx, ihy, ihz;  // 1 / (2h) per axis, for them
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

// Loads are unconditional: plane xx clamped into the grid (a uniform offset
// `po`) at a 32-bit lane byte offset `o` inside the plane (0 for lanes
// outside the grid). Values loaded for points outside the grid only ever feed boundary
// voxels, whose outputs are overwritten with zero, so no select follows the
// load and the load's latency stays hidden behind the next barrier. The
// address is a scalar base plus the lane offset: no 64-bit arithmetic per
// load. Planes hold fewer than kMaxMarchPlane voxels (checked at launch).
template <typename T, bool INCR>
__device__ __forceinline__ Col<T> load_col(const StepArgs<T>& a, int64_t po, uint32_t o) {
    const T u0 = at_byte(a.A[0] + po, o), u1 = at_byte(a.A[1] + po, o), u2 = at_byte(a.A[2] + po, o);
    const T b0 = at_byte(a.B[0] + po, o), b1 = at_byte(a.B[1] + po, o), b2 = at_byte(a.B[2] + po, o);
    Col<T> c;
    c.u0 = u0;
    c.u1 = u1;
    c.u2 = u2;
    c.d0 = INCR ? b0 : u0 - b0;
    c.d1 = INCR ? b1 : u1 - b1;
    c.d2 = INCR ? b2 : u2 - b2;
    return c;
}

// Coefficients and force are single-use streams (load_once: non-temporal).
template <typename T, bool VISC, bool FORCE>
__device__ __forceinline__ Coef<T> load_coef(const StepArgs<T>& a, int64_t po, uint32_t o) {
    Coef<T> k;
    k.m = load_once(&at_byte(a.m + po, o));
    k.e = VISC ? load_once(&at_byte(a.e + po, o)) : T(0);
    k.f0 = FORCE ? load_once(&at_byte(a.f[0] + po, o)) : T(0);
    k.f1 = FORCE ? load_once(&at_byte(a.f[1] + po, o)) : T(0);
    k.f2 = FORCE ? load_once(&at_byte(a.f[2] + po, o)) : T(0);
    return k;
}

// The same coefficients through the caches (plain loads), for the two-step
// pass: its ring points read the coefficients of the neighbouring tiles'
// edge columns, scattered accesses (one row per lane) that cost more from
// HBM than from the L2 (256^3 float32 pass: 219 -> 213 us).
template <typename T, bool VISC, bool FORCE>
__device__ __forceinline__ Coef<T> load_coef_cached(const StepArgs<T>& a, int64_t po, uint32_t o) {
    Coef<T> k;
    k.m = at_byte(a.m + po, o);
    k.e = VISC ? T(at_byte(a.e + po, o)) : T(0);
    k.f0 = FORCE ? T(at_byte(a.f[0] + po, o)) : T(0);
    k.f1 = FORCE ? T(at_byte(a.f[1] + po, o)) : T(0);
    k.f2 = FORCE ? T(at_byte(a.f[2] + po, o)) : T(0);
    return k;
}

// The next step's state is written with store_out (non-temporal): it is
// re-read one launch later, so the L2 is left to halo reuse.

// Increment of one component at one point (solver.py:194-234 in increment
// form): (d + m lap(u) + e lap(d) + f) (1 - vel_damp), with lap the 7-point
// Laplacian scaled per axis. Contraction is off and every fused multiply-add
// is written out, so the result does not depend on the kernel it is inlined
// into: the increments of a multi-step pass equal those of single steps bit
// for bit.
template <typename T, bool VISC, bool FORCE>
__device__ __forceinline__ T comp_inc(const StepArgs<T>& a, T u0, T um, T up, T uyp, T uym, T uzp, T uzm, T d0, T dm,
                                      T dp, T dyp, T dym, T dzp, T dzm, T m, T e, T f) {
#pragma clang fp contract(off)
    const T lu = fma(a.cx, fma(T(-2), u0, up + um), fma(a.cy, fma(T(-2), u0, uyp + uym), a.cz * fma(T(-2), u0, uzp + uzm)));
    T inc = fma(m, lu, d0);
    if (VISC) {
        const T ld =
            fma(a.vx, fma(T(-2), d0, dp + dm), fma(a.vy, fma(T(-2), d0, dyp + dym), a.vz * fma(T(-2), d0, dzp + dzm)));
        inc = fma(e, ld, inc);
    }
    if (FORCE && a.fs != T(0)) inc = inc + f;  // a step with env = 0 adds nothing, as the force-free kernel
    return inc * a.damp;
}

// Increments of the three components at one point: m, c, p the point's
// column at x-1, x, x+1; L an LDS plane of fields u0, u1, u2 (and d0, d1, d2
// with viscosity) holding the point at [r][q] and its y/z neighbours.
template <typename T, bool VISC, bool FORCE, typename Plane>
__device__ __forceinline__ void point_inc(const StepArgs<T>& a, const Col<T>& m, const Col<T>& c, const Col<T>& p,
                                          const Plane& L, int r, int q, const Coef<T>& k, T& i0, T& i1, T& i2) {
    constexpr int NL = VISC ? 6 : 3;
    auto nb = [&](int f, int dy, int dz) { return L[f][r + dy][q + dz]; };
    auto dn = [&](int f, int dy, int dz) { return VISC ? L[NL - 3 + f][r + dy][q + dz] : T(0); };
    i0 = comp_inc<T, VISC, FORCE>(a, c.u0, m.u0, p.u0, nb(0, 1, 0), nb(0, -1, 0), nb(0, 0, 1), nb(0, 0, -1), c.d0,
                                  m.d0, p.d0, dn(0, 1, 0), dn(0, -1, 0), dn(0, 0, 1), dn(0, 0, -1), k.m, k.e,
                                  FORCE ? a.fs * k.f0 : T(0));
    i1 = comp_inc<T, VISC, FORCE>(a, c.u1, m.u1, p.u1, nb(1, 1, 0), nb(1, -1, 0), nb(1, 0, 1), nb(1, 0, -1), c.d1,
                                  m.d1, p.d1, dn(1, 1, 0), dn(1, -1, 0), dn(1, 0, 1), dn(1, 0, -1), k.m, k.e,
                                  FORCE ? a.fs * k.f1 : T(0));
    i2 = comp_inc<T, VISC, FORCE>(a, c.u2, m.u2, p.u2, nb(2, 1, 0), nb(2, -1, 0), nb(2, 0, 1), nb(2, 0, -1), c.d2,
                                  m.d2, p.d2, dn(2, 1, 0), dn(2, -1, 0), dn(2, 0, 1), dn(2, 0, -1), k.m, k.e,
                                  FORCE ? a.fs * k.f2 : T(0));
}

// One tile of the time step: columns (y0 .. y0+BY-1, z0 .. z0+BZ-1), planes
// x0 .. x1-1; the block is (BZ, ROWS >= BY) threads (BZ a multiple of 64, so a
// wave never spans two rows) and all of them call this together; rows BY and
// up only keep the barriers. `lds` holds at least NL = (VISC ? 6 : 3) planes
// of (BY+2) x (BZ+2).
template <typename T, bool INCR, bool VISC, bool FORCE, int BY, int BZ, int NLD, int ROWS = BY>
__device__ __forceinline__ void march_tile(const StepArgs<T>& a, int x0, int x1, int y0, int z0,
                                           T (&lds)[2][NLD][BY + 2][BZ + 2]) {
    constexpr int NL = VISC ? 6 : 3;  // planes staged in LDS: u (3) and, with viscosity, d (3)
    static_assert(NLD >= NL, "LDS too small");
    static_assert(BZ % 64 == 0 && BY >= 2, "tile shape");
    const Grid g = a.g;
    const int tz = threadIdx.x, ty = threadIdx.y;

    const bool live = ROWS == BY || ty < BY;  // rows past BY only take part in the barriers
    const int y = y0 + ty, z = z0 + tz;
    const bool inyz = live && (y < g.ny) && (z < g.nz);
    const int64_t P = g.plane();
    const uint32_t yz = inyz ? (uint32_t)(y * g.nz + z) * (uint32_t)sizeof(T) : 0u;  // bytes
    const bool yz_interior = inyz && y > 0 && y < g.ny - 1 && z > 0 && z < g.nz - 1;

    // Halo of the tile. Row halos (y0-1, y0+BY) are loaded by the first and
    // last row of threads, at their own z (wave-uniform); column halos
    // (z0-1, z0+BZ) by the first and last lane of each row. A halo point
    // outside the grid is read at offset 0 of its plane and only feeds
    // boundary voxels.
    const bool row_thread = ty == 0 || ty == BY - 1;
    const int ry = ty == 0 ? y0 - 1 : y0 + BY;
    const int rly = ty == 0 ? 0 : BY + 1;
    const bool row_in = row_thread && ry >= 0 && ry < g.ny && z < g.nz;
    const uint32_t ryz = row_in ? (uint32_t)(ry * g.nz + z) * (uint32_t)sizeof(T) : 0u;
    const bool col_thread = live && (tz == 0 || tz == BZ - 1);
    const int cz = tz == 0 ? z0 - 1 : z0 + BZ;
    const int clz = tz == 0 ? 0 : BZ + 1;
    const bool col_in = col_thread && cz >= 0 && cz < g.nz && y < g.ny;
    const uint32_t cyz = col_in ? (uint32_t)(y * g.nz + cz) * (uint32_t)sizeof(T) : 0u;

    auto po = [&](int xx) { return (int64_t)min(max(xx, 0), g.nx - 1) * P; };
    auto center = [&](int xx) { return load_col<T, INCR>(a, po(xx), yz); };
    auto row_halo = [&](int xx) {
        Col<T> c{};
        if (row_thread) c = load_col<T, INCR>(a, po(xx), ryz);
        return c;
    };
    auto col_halo = [&](int xx) {
        Col<T> c{};
        if (col_thread) c = load_col<T, INCR>(a, po(xx), cyz);
        return c;
    };
    auto coef = [&](int xx) { return load_coef<T, VISC, FORCE>(a, po(xx), yz); };

    // One x plane. m, c, p: planes x-1, x, x+1 of this column; n receives x+2.
    // hc: this plane's halo point, hn receives the next one; kc / kn likewise for coefficients.
    auto plane = [&](auto buf, const Col<T>& m, const Col<T>& c, const Col<T>& p, Col<T>& n, const Col<T>& hc,
                     Col<T>& hn, const Col<T>& vc, Col<T>& vn, const Coef<T>& kc, Coef<T>& kn, int x) {
        constexpr int B = decltype(buf)::value;
        auto& L = lds[B];
        if (live) {
            L[0][ty + 1][tz + 1] = c.u0;
            L[1][ty + 1][tz + 1] = c.u1;
            L[2][ty + 1][tz + 1] = c.u2;
            if (VISC) {
                L[NL - 3][ty + 1][tz + 1] = c.d0;
                L[NL - 2][ty + 1][tz + 1] = c.d1;
                L[NL - 1][ty + 1][tz + 1] = c.d2;
            }
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
        T i0, i1, i2;
        point_inc<T, VISC, FORCE>(a, m, c, p, L, ty + 1, tz + 1, kc, i0, i1, i2);
        if (!in) i0 = i1 = i2 = T(0);
        const int64_t px = (int64_t)x * P;
        store_out(&at_byte(a.Ao[0] + px, yz), in ? c.u0 + i0 : T(0));
        store_out(&at_byte(a.Ao[1] + px, yz), in ? c.u1 + i1 : T(0));
        store_out(&at_byte(a.Ao[2] + px, yz), in ? c.u2 + i2 : T(0));
        if (INCR) {
            store_out(&at_byte(a.Bo[0] + px, yz), i0);
            store_out(&at_byte(a.Bo[1] + px, yz), i1);
            store_out(&at_byte(a.Bo[2] + px, yz), i2);
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
