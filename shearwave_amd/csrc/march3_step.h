// Three time steps per pass over the state (deeper temporal blocking of the
// loop in shear_fdtd, solver.py:194-234): one x-marching tile reads u_n, d_n
// and the coefficients once and writes u_{n+3}, d_{n+3}. Force-free steps,
// increment form.
//
// Tile: output columns (y0 .. y0+BY-1, z0 .. z0+63). Step n+3 at a point
// needs step n+2 at its six neighbours, step n+2 needs step n+1, so step n+2
// is evaluated on the tile plus a one-voxel ring, step n+1 on the tile plus
// two voxels, from step n on the tile plus three. The block is 64 x (BY + 5)
// threads, one wave per row:
//   rows 0 .. BY-1   output rows y0+ty: steps n+1, n+2, n+3;
//   rows BY, BY+1    ring-1 rows y0-1, y0+BY: steps n+1, n+2;
//   rows BY+2, BY+3  ring-2 rows y0-2, y0+BY+1: step n+1; they also load
//                    step n of the outer rows y0-3, y0+BY+2;
//   row  BY+4        the ring columns, lane j < 4(BY+4): column z0-1, z0+64,
//                    z0-2 or z0+65 (j / (BY+4)) at row y0-2 + j % (BY+4):
//                    step n+1, and step n+2 at the ring-1 points; lanes
//                    below 2(BY+6)+8 also load one outer step-n point
//                    (columns z0-3 and z0+66, and the eight corners of rows
//                    y0-3, y0+BY+2).
// Each x iteration stages step n of plane x in LDS (tile + 3), evaluates
// step n+1 at plane x (tile + 2, staged in a second LDS plane), step n+2 at
// plane x-1 (tile + 1, staged in a third) and step n+3 at plane x-2. All
// three LDS planes are double-buffered, so one barrier per plane suffices.
// Every step calls the single-step arithmetic (point_inc), so the increments
// equal those of three single steps bit for bit; the stored displacement
// combines the last two increments as the two-step pass does,
// u_{n+3} = u_{n+1} + (d_{n+2} + d_{n+3}).
#pragma once

#include <type_traits>

#include "march2_step.h"
#include "march_step.h"

namespace sw {

template <typename T, bool VISC, int BY>
struct Step3Lds {
    static constexpr int NF = VISC ? 6 : 3;  // u and, with viscosity, d
    T s0[2][NF][BY + 6][70];                 // step n, three-voxel halo
    T s1[2][NF][BY + 4][68];                 // step n+1, two-voxel halo
    T s2[2][NF][BY + 2][66];                 // step n+2, one-voxel halo
};

template <typename T, bool VISC, int BY>
__device__ __forceinline__ void march3_tile(const StepArgs<T>& a, int x0, int x1, int y0, int z0,
                                            Step3Lds<T, VISC, BY>& lds) {
    static_assert(4 * (BY + 4) <= 64 && 2 * (BY + 6) + 8 <= 64 && BY + 5 <= 16, "tile rows");
    constexpr bool FORCE = false;
    const Grid g = a.g;
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int64_t P = g.plane();

    // this thread's step-n+1 point (prim) and outer step-n point (outer);
    // lvl: the last step evaluated at the prim point (1, 2 or 3)
    int py = y0, pz = z0, oy = y0, oz = z0, lvl = 1;
    bool prim = true, outer = false;
    if (ty < BY) {
        py = y0 + ty;
        pz = z0 + tz;
        lvl = 3;
    } else if (ty < BY + 2) {
        py = ty == BY ? y0 - 1 : y0 + BY;
        pz = z0 + tz;
        lvl = 2;
    } else if (ty < BY + 4) {
        py = ty == BY + 2 ? y0 - 2 : y0 + BY + 1;
        pz = z0 + tz;
        oy = ty == BY + 2 ? y0 - 3 : y0 + BY + 2;
        oz = pz;
        outer = true;
    } else {
        constexpr int R = BY + 4;  // ring rows y0-2 .. y0+BY+1
        if (tz < 4 * R) {
            const int c = tz / R, j = tz % R;
            py = y0 - 2 + j;
            pz = c == 0 ? z0 - 1 : c == 1 ? z0 + 64 : c == 2 ? z0 - 2 : z0 + 65;
            lvl = (c < 2 && j >= 1 && j <= BY + 2) ? 2 : 1;
        } else {
            prim = false;
        }
        constexpr int RO = BY + 6;  // outer rows y0-3 .. y0+BY+2
        if (tz < 2 * RO) {
            oy = y0 - 3 + tz % RO;
            oz = tz < RO ? z0 - 3 : z0 + 66;
            outer = true;
        } else if (tz < 2 * RO + 8) {
            const int k = tz - 2 * RO;
            oy = k < 4 ? y0 - 3 : y0 + BY + 2;
            const int c = k & 3;
            oz = c == 0 ? z0 - 2 : c == 1 ? z0 - 1 : c == 2 ? z0 + 64 : z0 + 65;
            outer = true;
        }
    }
    auto in_grid = [&](int y, int z) { return y >= 0 && y < g.ny && z >= 0 && z < g.nz; };
    const bool p_in = prim && in_grid(py, pz);
    const bool p_int = p_in && py > 0 && py < g.ny - 1 && pz > 0 && pz < g.nz - 1;
    const uint32_t pyz = p_in ? (uint32_t)(py * g.nz + pz) * (uint32_t)sizeof(T) : 0u;
    const bool o_in = outer && in_grid(oy, oz);
    const uint32_t oyz = o_in ? (uint32_t)(oy * g.nz + oz) * (uint32_t)sizeof(T) : 0u;
    const int r0 = py - y0 + 3, q0 = pz - z0 + 3;  // in the step-n plane
    const int r1 = py - y0 + 2, q1 = pz - z0 + 2;  // in the step-n+1 plane
    const int r2 = py - y0 + 1, q2 = pz - z0 + 1;  // in the step-n+2 plane
    const int ro = oy - y0 + 3, qo = oz - z0 + 3;
    const bool do2 = prim && lvl >= 2, do3 = prim && lvl == 3;

    auto po = [&](int xx) { return (int64_t)min(max(xx, 0), g.nx - 1) * P; };
    auto col = [&](int xx, uint32_t o) { return load_col<T, true>(a, po(xx), o); };
    auto coef = [&](int xx) { return load_coef_cached<T, VISC, FORCE>(a, po(xx), pyz); };
    auto outer_col = [&](int xx) {
        Col<T> c{};
        if (outer) c = col(xx, oyz);
        return c;
    };
    // a step's result at a point: u + inc and inc, or zero off the interior
    auto next = [&](const Col<T>& c, bool in, T i0, T i1, T i2) {
        Col<T> n;
        n.u0 = in ? c.u0 + i0 : T(0);
        n.u1 = in ? c.u1 + i1 : T(0);
        n.u2 = in ? c.u2 + i2 : T(0);
        n.d0 = in ? i0 : T(0);
        n.d1 = in ? i1 : T(0);
        n.d2 = in ? i2 : T(0);
        return n;
    };

    // rings indexed by plane: slot (plane - xb) & 3; one plane of prefetch
    Col<T> q[4];   // step n of this column: planes x-1 .. x+2
    Coef<T> k[4];  // coefficients: planes x-2 .. x+1
    Col<T> o[4];   // step n at the outer point: planes x .. x+1
    Col<T> s1[4];  // step n+1 of this column: planes x-2 .. x
    Col<T> s2[4];  // step n+2 of this column: planes x-3 .. x-1
    const int xb = x0 - 2, xe = x1 + 1;
#pragma unroll
    for (int j = -1; j <= 1; ++j) q[(j + 4) & 3] = col(xb + j, pyz);  // planes xb-1 .. xb+1
    k[0] = coef(xb);
    o[0] = outer_col(xb);

    march_incl<4>(xb, xe, [&](auto kc, int x) {
        constexpr int K = decltype(kc)::value;
        constexpr int B = K & 1;
        auto& S0 = lds.s0[B];
        // stage step n of plane x
        if (prim) stage<T, VISC>(S0, r0, q0, q[K]);
        if (outer) stage<T, VISC>(S0, ro, qo, o[K]);
        // prefetch: step n at x+2, coefficients and outer point at x+1
        q[(K + 2) & 3] = col(x + 2, pyz);
        k[(K + 1) & 3] = coef(x + 1);
        o[(K + 1) & 3] = outer_col(x + 1);
        __syncthreads();
        // step n+1 at plane x on the tile + 2
        if (prim) {
            T i0, i1, i2;
            point_inc<T, VISC, FORCE>(a, q[(K + 3) & 3], q[K], q[(K + 1) & 3], S0, r0, q0, k[K], i0, i1, i2);
            const Col<T> n = next(q[K], p_int && x > 0 && x < g.nx - 1, i0, i1, i2);
            s1[K] = n;
            stage<T, VISC>(lds.s1[B], r1, q1, n);
        }
        // step n+2 at plane x-1 on the tile + 1 (step n+1 of plane x-1 was staged one plane ago)
        if (do2 && x - 1 >= x0 - 1) {
            const Col<T>& c = s1[(K + 3) & 3];
            T i0, i1, i2;
            point_inc<T, VISC, FORCE>(a, s1[(K + 2) & 3], c, s1[K], lds.s1[B ^ 1], r1, q1, k[(K + 3) & 3], i0, i1,
                                      i2);
            const Col<T> n = next(c, p_int && x - 1 > 0 && x - 1 < g.nx - 1, i0, i1, i2);
            s2[(K + 3) & 3] = n;
            stage<T, VISC>(lds.s2[B ^ 1], r2, q2, n);
        }
        // step n+3 at plane x-2 on the tile (step n+2 of plane x-2 was staged one plane ago)
        if (do3 && x - 2 >= x0 && p_in) {
            const Col<T>& c = s2[(K + 2) & 3];
            T i0, i1, i2;
            point_inc<T, VISC, FORCE>(a, s2[(K + 1) & 3], c, s2[(K + 3) & 3], lds.s2[B], r2, q2, k[(K + 2) & 3], i0,
                                      i1, i2);
            const bool in = p_int && x - 2 > 0 && x - 2 < g.nx - 1;
            if (!in) i0 = i1 = i2 = T(0);
            const int64_t px = (int64_t)(x - 2) * P;
            const Col<T>& u1 = s1[(K + 2) & 3];  // step n+1 of plane x-2
            store_out(&at_byte(a.Ao[0] + px, pyz), in ? u1.u0 + (c.d0 + i0) : T(0));
            store_out(&at_byte(a.Ao[1] + px, pyz), in ? u1.u1 + (c.d1 + i1) : T(0));
            store_out(&at_byte(a.Ao[2] + px, pyz), in ? u1.u2 + (c.d2 + i2) : T(0));
            store_out(&at_byte(a.Bo[0] + px, pyz), i0);
            store_out(&at_byte(a.Bo[1] + px, pyz), i1);
            store_out(&at_byte(a.Bo[2] + px, pyz), i2);
        }
    });
}

}  // namespace sw
