// Two time steps per pass over the state (temporal blocking of the loop in
// shear_fdtd, solver.py:194-234): one x-marching tile reads u_n, d_n, the
// coefficients (and the force, when either step applies it) once and writes
// u_{n+2}, d_{n+2}, halving the HBM traffic per voxel-step of the single-step
// tile. Increment form; step n+1 scales the force by a.fs, step n+2 by a.fs2.
//
// Tile: output columns (y0 .. y0+BY-1, z0 .. z0+63). Step n+2 at a point
// needs step n+1 at its six neighbours, so step n+1 is evaluated on the tile
// plus a one-voxel ring, which needs step n on the tile plus two voxels.
// The block is 64 x (BY + 3) threads (one wave per row):
//   rows 0 .. BY-1   output points: step n+1 and step n+2;
//   row  BY          ring row y0-1 (step n+1), loads outer row y0-2;
//   row  BY+1        ring row y0+BY (step n+1), loads outer row y0+BY+1;
//   row  BY+2        lanes 0..BY-1: ring column z0-1 at row y0+lane, loads z0-2;
//                    lanes BY..2BY-1: ring column z0+64, loads z0+65;
//                    lanes 2BY..2BY+3: the four ring corners (step n only).
// Each x iteration stages step n of plane x in LDS (tile + 2), evaluates
// step n+1 at plane x on the tile + 1 ring (staged in a second LDS plane),
// and step n+2 at plane x-1 from the step n+1 planes x-2, x-1, x. Both LDS
// planes are double-buffered, so one barrier per plane suffices. Step n+1
// and n+2 call the single-step arithmetic (point_inc), so both increments
// equal those of two single steps bit for bit. The stored displacement takes
// one rounding per pass, u_{n+2} = u_n + (d_{n+1} + d_{n+2}), where two single
// steps round twice, (u_n + d_{n+1}) + d_{n+2}: late in a solve |d| is 1e-3 of
// |u| or less, and each rounding of u in float32 loses a visible part of the
// increment (on the 256^3, 2000-step workload this halves the roundings of u
// and takes the float32 error against float64 from 4.8e-6 to 1.2e-6 on u and
// from 1.2e-5 to 3.4e-6 on v).
#pragma once

#include <type_traits>

#include "march_step.h"
#include "strain_point.h"

namespace sw {

template <typename T, bool VISC, int BY>
struct Step2Lds {
    static constexpr int NF = VISC ? 6 : 3;  // u and, with viscosity, d
    T s0[2][NF][BY + 4][68];                 // step n, two-voxel halo
    T s1[2][NF][BY + 2][66];                 // step n+1, one-voxel halo
};

template <int K>
using Kc = std::integral_constant<int, K>;

// body(Kc<k>{}, x) for x = xb .. xe (inclusive) with k = (x - xb) mod U, U = 4 or 8.
template <int U, typename F>
__device__ __forceinline__ void march_incl(int xb, int xe, F&& body) {
    static_assert(U == 4 || U == 8, "unroll");
    for (int x = xb; x <= xe; x += U) {
        body(Kc<0>{}, x);
        if (x + 1 > xe) break;
        body(Kc<1>{}, x + 1);
        if (x + 2 > xe) break;
        body(Kc<2>{}, x + 2);
        if (x + 3 > xe) break;
        body(Kc<3>{}, x + 3);
        if constexpr (U == 8) {
            if (x + 4 > xe) break;
            body(Kc<4 % U>{}, x + 4);
            if (x + 5 > xe) break;
            body(Kc<5 % U>{}, x + 5);
            if (x + 6 > xe) break;
            body(Kc<6 % U>{}, x + 6);
            if (x + 7 > xe) break;
            body(Kc<7 % U>{}, x + 7);
        }
    }
}

// Slot of a ring of N entries for the plane `off` planes from the current
// one, in iteration K of a loop unrolled by a multiple of N.
template <int N>
constexpr int ring(int k) {
    return ((k % N) + N) % N;
}

// Planes of step n loaded ahead of the one being staged (1 or 2; the
// coefficient ring holds x-1 .. x+2). A block holds one CU on its own, so its
// waves must keep several planes of loads in flight; two planes fit the
// registers of the float32 8-row tile (11 waves), one plane the others.
template <typename T, int BY, bool FORCE>
constexpr int step2_dist() {
#ifdef SW_STEP2_DIST
    return SW_STEP2_DIST;
#else
    return (sizeof(T) == 4 && BY <= 8 && !FORCE) ? 2 : 1;
#endif
}

// Write a column's fields into plane L at [r][q].
template <typename T, bool VISC, typename Plane>
__device__ __forceinline__ void stage(Plane& L, int r, int q, const Col<T>& c) {
    L[0][r][q] = c.u0;
    L[1][r][q] = c.u1;
    L[2][r][q] = c.u2;
    if (VISC) {
        L[3][r][q] = c.d0;
        L[4][r][q] = c.d1;
        L[5][r][q] = c.d2;
    }
}

// With MID the pass also stores step n+1 of its own planes (u_{n+1} = u_n +
// d_{n+1} and d_{n+1}, the values two single steps store) into a.Am, a.Bm.
// With STRAIN it also stores the in-loop strain invariants of its input
// state u_n into a.inv (strain_point.h): u_n's x neighbours are in the
// column's queue and its y/z neighbours in the step-n LDS plane, so the
// invariants cost their 16 B of stores per voxel and no read.
// DIST overrides the prefetch distance (0: step2_dist).
template <typename T, bool VISC, bool FORCE, int BY, bool MID = false, int DIST = 0, bool STRAIN = false>
__device__ __forceinline__ void march2_tile(const StepArgs<T>& a, int x0, int x1, int y0, int z0,
                                            Step2Lds<T, VISC, BY>& lds) {
    static_assert(2 * BY + 4 <= 64 && BY + 3 <= 16, "tile rows");
    const Grid g = a.g;
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int64_t P = g.plane();

    // this thread's step-n+1 point (prim) and outer step-n point (outer)
    int py = y0, pz = z0, oy = y0, oz = z0;
    bool prim = true, outer = false;
    if (ty < BY) {
        py = y0 + ty;
        pz = z0 + tz;
    } else if (ty == BY || ty == BY + 1) {
        py = ty == BY ? y0 - 1 : y0 + BY;
        pz = z0 + tz;
        oy = ty == BY ? y0 - 2 : y0 + BY + 1;
        oz = pz;
        outer = true;
    } else if (tz < 2 * BY) {
        py = oy = y0 + (tz < BY ? tz : tz - BY);
        pz = tz < BY ? z0 - 1 : z0 + 64;
        oz = tz < BY ? z0 - 2 : z0 + 65;
        outer = true;
    } else if (tz < 2 * BY + 4) {
        const int k = tz - 2 * BY;
        prim = false;
        oy = k < 2 ? y0 - 1 : y0 + BY;
        oz = (k & 1) ? z0 + 64 : z0 - 1;
        outer = true;
    } else {
        prim = false;
    }
    const bool out_thread = ty < BY;
    auto in_grid = [&](int y, int z) { return y >= 0 && y < g.ny && z >= 0 && z < g.nz; };
    const bool p_in = prim && in_grid(py, pz);
    const bool p_int = p_in && py > 0 && py < g.ny - 1 && pz > 0 && pz < g.nz - 1;
    const uint32_t pyz = p_in ? (uint32_t)(py * g.nz + pz) * (uint32_t)sizeof(T) : 0u;
    const bool o_in = outer && in_grid(oy, oz);
    const uint32_t oyz = o_in ? (uint32_t)(oy * g.nz + oz) * (uint32_t)sizeof(T) : 0u;
    const int r0 = py - y0 + 2, q0 = pz - z0 + 2;  // in the step-n plane
    const int r1 = py - y0 + 1, q1 = pz - z0 + 1;  // in the step-n+1 plane
    const int ro = oy - y0 + 2, qo = oz - z0 + 2;

    auto po = [&](int xx) { return (int64_t)min(max(xx, 0), g.nx - 1) * P; };
    auto col = [&](int xx, uint32_t o) { return load_col<T, true>(a, po(xx), o); };
    auto coef = [&](int xx) { return load_coef_cached<T, VISC, FORCE>(a, po(xx), pyz); };
    StepArgs<T> a2 = a;  // step n+2: the second step's force scale
    a2.fs = a.fs2;
    auto outer_col = [&](int xx) {
        Col<T> c{};
        if (outer) c = col(xx, oyz);
        return c;
    };

    // rings: a plane p sits in slot (p - xb) mod (ring size), xb = x0 - 1 the first
    // iteration; the loop is unrolled by QN, a multiple of every ring size, so
    // each slot is a compile-time index
    // (a third plane of prefetch, with rings of 6, measured no faster at 256^3)
    constexpr int D = DIST > 0 ? DIST : step2_dist<T, BY, FORCE>();
    static_assert(D >= 1 && D <= 2, "two-step prefetch distance");
    constexpr int QN = D == 1 ? 4 : 8;  // unroll; ring of q: planes x-1 .. x+1+D
    constexpr int KN = 4;               // ring of k (planes x-1 .. x+D) and o (x .. x+D)
    constexpr int SN = 4;               // ring of s: planes x-2 .. x
    static_assert(QN >= D + 3 && KN >= D + 2 && SN >= 3 && QN % KN == 0 && QN % SN == 0 && QN % 2 == 0, "rings");
    Col<T> q[QN];   // step n of this column: planes x-1 .. x+1+D
    Coef<T> k[KN];  // coefficients: planes x-1 .. x+D
    Col<T> s[SN];   // step n+1 of this column: planes x-2 .. x
    Col<T> o[KN];   // step n at the outer point: planes x .. x+D
    const int xb = x0 - 1;
#pragma unroll
    for (int j = -1; j <= D; ++j) q[ring<QN>(j)] = col(x0 - 1 + j, pyz);  // planes x0-2 .. x0-1+D
#pragma unroll
    for (int j = 0; j < D; ++j) {
        k[j] = coef(x0 - 1 + j);
        o[j] = outer_col(x0 - 1 + j);
    }

    march_incl<QN>(xb, x1, [&](auto kc, int x) {
        constexpr int K = decltype(kc)::value;
        constexpr int B = K & 1;
        auto& S0 = lds.s0[B];
        auto& S1 = lds.s1[B];
        // stage step n of plane x
        if (prim) stage<T, VISC>(S0, r0, q0, q[K]);
        if (outer) stage<T, VISC>(S0, ro, qo, o[ring<KN>(K)]);
        // prefetch: step n at x+1+D, coefficients and outer point at x+D
        q[ring<QN>(K + 1 + D)] = col(x + 1 + D, pyz);
        k[ring<KN>(K + D)] = coef(x + D);
        o[ring<KN>(K + D)] = outer_col(x + D);
        __syncthreads();
        if constexpr (STRAIN) {  // the invariants of step n at plane x
            if (out_thread && p_in && x >= x0 && x < x1) {
                const Col<T>& m = q[ring<QN>(K - 1)];
                const Col<T>& p = q[ring<QN>(K + 1)];
                auto dy = [&](int f) { return S0[f][r0 + 1][q0] - S0[f][r0 - 1][q0]; };
                auto dz = [&](int f) { return S0[f][r0][q0 + 1] - S0[f][r0][q0 - 1]; };
                const UDiff<T> du{p.u0 - m.u0, p.u1 - m.u1, p.u2 - m.u2, dy(0), dy(1), dy(2), dz(0), dz(1), dz(2)};
                T vol, dev, ms, vm;
                inloop_invariants(p_int && x > 0 && x < g.nx - 1, du, a.ihx, a.ihy, a.ihz, vol, dev, ms, vm);
                const int64_t px = (int64_t)x * P;
                store_out(&at_byte(a.inv[0] + px, pyz), vol);
                store_out(&at_byte(a.inv[1] + px, pyz), dev);
                store_out(&at_byte(a.inv[2] + px, pyz), ms);
                store_out(&at_byte(a.inv[3] + px, pyz), vm);
            }
        }
        // step n+1 at plane x on the tile and its ring
        {
            T i0, i1, i2;
            point_inc<T, VISC, FORCE>(a, q[ring<QN>(K - 1)], q[K], q[ring<QN>(K + 1)], S0, r0, q0, k[ring<KN>(K)], i0,
                                      i1, i2);
            const bool in = p_int && x > 0 && x < g.nx - 1;
            Col<T> n;
            n.u0 = in ? q[K].u0 + i0 : T(0);
            n.u1 = in ? q[K].u1 + i1 : T(0);
            n.u2 = in ? q[K].u2 + i2 : T(0);
            n.d0 = in ? i0 : T(0);
            n.d1 = in ? i1 : T(0);
            n.d2 = in ? i2 : T(0);
            s[ring<SN>(K)] = n;
            if (prim) stage<T, VISC>(S1, r1, q1, n);
            if constexpr (MID) {
                if (out_thread && p_in && x >= x0 && x < x1) {
                    const int64_t pxm = (int64_t)x * P;
                    store_out(&at_byte(a.Am[0] + pxm, pyz), n.u0);
                    store_out(&at_byte(a.Am[1] + pxm, pyz), n.u1);
                    store_out(&at_byte(a.Am[2] + pxm, pyz), n.u2);
                    store_out(&at_byte(a.Bm[0] + pxm, pyz), n.d0);
                    store_out(&at_byte(a.Bm[1] + pxm, pyz), n.d1);
                    store_out(&at_byte(a.Bm[2] + pxm, pyz), n.d2);
                }
            }
        }
        // step n+2 at plane x-1 (staged in the other step-n+1 buffer one plane ago)
        if (out_thread && x - 1 >= x0 && p_in) {
            auto& S1p = lds.s1[B ^ 1];
            const Col<T>& c = s[ring<SN>(K - 1)];
            T i0, i1, i2;
            point_inc<T, VISC, FORCE>(a2, s[ring<SN>(K - 2)], c, s[ring<SN>(K)], S1p, r1, q1, k[ring<KN>(K - 1)], i0,
                                      i1, i2);
            const bool in = p_int && x - 1 > 0 && x - 1 < g.nx - 1;
            if (!in) i0 = i1 = i2 = T(0);
            const int64_t px = (int64_t)(x - 1) * P;
            const Col<T>& un = q[ring<QN>(K - 1)];  // step n of plane x-1
            store_out(&at_byte(a.Ao[0] + px, pyz), in ? un.u0 + (c.d0 + i0) : T(0));
            store_out(&at_byte(a.Ao[1] + px, pyz), in ? un.u1 + (c.d1 + i1) : T(0));
            store_out(&at_byte(a.Ao[2] + px, pyz), in ? un.u2 + (c.d2 + i2) : T(0));
            store_out(&at_byte(a.Bo[0] + px, pyz), i0);
            store_out(&at_byte(a.Bo[1] + px, pyz), i1);
            store_out(&at_byte(a.Bo[2] + px, pyz), i2);
        }
    });
}

}  // namespace sw
