// Shared device code of the projection's conjugate gradients (poisson_cg,
// solver.py:1010-1050): the device-side CG state, deterministic slot
// reductions, and the x-marching tiles of the two fused kernels of one
// iteration. Used by the single-solve kernels (project.hip) and the batched
// ones (batch.hip), so a solve's arithmetic is identical in both.
#pragma once

#include <cmath>
#include <type_traits>
#include <utility>

#include "sw_common.h"

namespace sw {

// Device-resident state of one CG solve. The host only reads stop/reason/iters.
struct CgState {
    double rs_old, rs0, alpha, beta, relres, max_div, tol;
    double alpha_prev;  // alpha of the iteration before (deferred phi updates)
    long long iters;
    int stop, reason;
    unsigned int counter;  // slots published in the current kernel
    int pad1;
    long long max_iters;   // stop after this many phi updates (solver.py:1033)
    double true_relres;    // sqrt(|b - A phi|^2 / rs0) from the last true-residual pass
    double final_tol;      // the caller's tol; `tol` may be a looser stage before a residual replacement
    double rz;             // r.z of the preconditioned CG (0 before the first application)
    int precond;           // 1: preconditioned (z = M r; alpha and beta from r.z, the stop from r.r)
    int pad3;
};
static_assert(sizeof(CgState) <= 128, "CgState too large");

__device__ inline double load_coherent(const double* p) {
    unsigned long long v =
        __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}

// Publish a block-reduced partial into slot `slot` of `n`; returns true (in
// every thread) in the block that published the last slot. The counter is
// reset by that block.
__device__ inline bool publish_slot(double* part_sum, double* part_max, int slot, double s, double m,
                                    unsigned int* counter, unsigned int n) {
    __shared__ int last;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if (tid == 0) {
        part_sum[slot] = s;
        if (part_max) part_max[slot] = m;
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = (prev == n - 1);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// Count k slots, stored by thread 0 of this block since its last count, as
// published (one fence and one atomic per work item); returns true (in every
// thread) in the block that completed the n slots.
__device__ inline bool publish_count(unsigned int* counter, unsigned int k, unsigned int n) {
    __shared__ int last;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if (tid == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(counter, k);
        last = (prev + k == n);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// In the last block: sum (and max) over the n slots in a fixed order.
__device__ inline void reduce_slots(const double* part_sum, const double* part_max, int n, double* red,
                                    double* out_sum, double* out_max) {
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nt = blockDim.x * blockDim.y;
    double s = 0.0, m = 0.0;
    for (int k = tid; k < n; k += nt) {
        s += load_coherent(part_sum + k);
        if (part_max) m = fmax(m, load_coherent(part_max + k));
    }
    *out_sum = block_sum(s, red);
    if (part_max) *out_max = block_max(m, red);
}

constexpr double kEps64 = 2.220446049250313e-16;

// After the divergence kernel: r0.r0 and max|div| (solver.py:672, 1028-1031).
__device__ inline void cg_after_init(CgState* st, double rs, double maxdiv) {
    st->rs_old = rs;
    st->rs0 = fmax(rs, kEps64);
    st->relres = sqrt(rs / fmax(rs, kEps64));
    st->max_div = maxdiv;
    st->counter = 0u;
    if (maxdiv < 1e-12) {  // nothing to remove
        st->stop = 1;
        st->reason = SW_CG_EARLY_EXIT;
        st->relres = 0.0;
    }
}

// After p.Ap (solver.py:1034-1038).
__device__ inline void cg_after_dir(CgState* st, double denom) {
    st->counter = 0u;
    if (!isfinite(denom) || denom <= 0.0) {
        st->stop = 1;
        st->reason = SW_CG_DENOM;
    } else {
        st->alpha_prev = st->alpha;
        st->alpha = (st->precond ? st->rz : st->rs_old) / denom;
    }
}

// After z = M r and r.z of the preconditioned CG: beta = r.z / previous r.z
// (0 for the first application, after which p = z).
__device__ inline void cg_after_precond(CgState* st, double rz) {
    st->counter = 0u;
    st->beta = st->rz > 0.0 ? rz / st->rz : 0.0;
    st->rz = rz;
    if (!isfinite(rz)) {
        st->stop = 1;
        st->reason = SW_CG_NONFINITE;
    }
}

// After r.r (solver.py:1041-1048).
__device__ inline void cg_after_step(CgState* st, double rs_new) {
    st->counter = 0u;
    st->iters += 1;
    if (!isfinite(rs_new)) {
        st->stop = 1;
        st->reason = SW_CG_NONFINITE;
        return;
    }
    const double rel = sqrt(rs_new / st->rs0);
    st->relres = rel;
    if (rel < st->tol) {
        st->stop = 1;
        st->reason = SW_CG_CONVERGED;
    } else {
        if (!st->precond) st->beta = rs_new / fmax(st->rs_old, kEps64);  // else from r.z after z = M r
        st->rs_old = rs_new;
        if (st->iters >= st->max_iters) {
            st->stop = 1;
            st->reason = SW_CG_MAX_ITERS;
        }
    }
}

// After a true-residual pass (r = b - A phi, r.r = rs). Without `replace` it
// only records the norm. With it, r replaces the recursively updated residual:
// the solve stops if the true norm meets the caller's tol (or the iteration
// budget is spent), and otherwise continues from it as from an ordinary
// iteration (beta = rs / rs_old, solver.py:1047) towards the stage tol `next_tol`.
__device__ inline void cg_after_true_residual(CgState* st, double rs, int replace, double next_tol) {
    st->counter = 0u;
    const double rel = sqrt(rs / st->rs0);
    st->true_relres = rel;
    if (!replace) return;
    st->relres = rel;
    if (!isfinite(rs)) {
        st->stop = 1;
        st->reason = SW_CG_NONFINITE;
    } else if (rel < st->final_tol) {
        st->stop = 1;
        st->reason = SW_CG_CONVERGED;
    } else if (st->iters >= st->max_iters) {
        st->stop = 1;
        st->reason = SW_CG_MAX_ITERS;
    } else {
        if (!st->precond) st->beta = rs / fmax(st->rs_old, kEps64);  // else from r.z after z = M r
        st->rs_old = rs;
        st->tol = next_tol;
        st->stop = 0;
        st->reason = SW_CG_RUNNING;
    }
}

// Fresh state for a solve with tolerance tol and at most max_iters iterations.
// `stage_tol` (>= tol) is where the recursion first stops for a true-residual
// check; pass tol when no residual replacement is planned.
__device__ inline void cg_reset(CgState* st, double tol, long long max_iters, double stage_tol) {
    st->rs_old = st->rs0 = st->alpha = st->beta = st->relres = st->max_div = st->alpha_prev = 0.0;
    st->true_relres = -1.0;
    st->rz = 0.0;
    st->precond = 0;
    st->pad3 = 0;
    st->final_tol = tol;
    st->tol = stage_tol;
    st->iters = 0;
    st->stop = max_iters <= 0 ? 1 : 0;
    st->reason = max_iters <= 0 ? SW_CG_MAX_ITERS : SW_CG_RUNNING;
    st->counter = 0u;
    st->max_iters = max_iters;
}

// ---------------------------------------------------------------------------
// x-marching CG tiles. A (BY x 64) tile of (y, z) columns marches along x; p
// at planes x-2 .. x+3 of the column sits in a register queue, and plane x
// (with a two-voxel halo in y and z, the reach of A) is staged in LDS, so r,
// p_old, p and phi are each read from HBM about once per iteration. The
// queue's roles rotate over a loop unrolled by six, so no prefetched register
// is moved.
//
// Vector storage V: r, p_old and p are double (as poisson_cg keeps them) or
// float; phi, every product, the dot products and the residual norm are
// double in both. With float vectors, p is rounded to float as it is formed
// and A p, p.Ap, phi += alpha p and r -= alpha A p all use that stored value,
// so phi and r stay consistent (b - A phi - r at double rounding level) and
// only the search directions carry float rounding; r is rounded as stored and
// r.r is summed over the stored values.

// v rounded to the vector storage type V and widened back (the identity for double).
template <typename V>
__device__ __forceinline__ double as_stored(double v) {
    return (double)(V)v;
}

#ifndef SW_CG_BY
#define SW_CG_BY 16
#endif
constexpr int kCgBY = SW_CG_BY;  // tile rows in y (block of 64 x kCgBY threads)
constexpr int kCgLZ = 68;  // 64 + 2 + 2

template <typename V>
using CgLdsT = V[2][kCgBY + 4][kCgLZ];

// Offsets inside a plane are 32-bit (planes of fewer than kMaxMarchPlane
// voxels, which the launchers check): a load is then a uniform plane pointer
// plus a lane offset, with no 64-bit address arithmetic per load.
struct CgTile {
    int x0, x1, y, z;
    bool inyz;
    uint32_t yz;   // element offset of (y, z) in a plane
    uint32_t yzb;  // the same in bytes of a double (vb<V>() gives it for V)
    bool halo_thread, halo_in;
    int ly, lz;
    uint32_t hyzb;  // byte offset of this thread's halo point (double)
};

// A double's byte offset b8 as the byte offset of the same element of a V array.
template <typename V>
__device__ __forceinline__ uint32_t vb(uint32_t b8) {
    return sizeof(V) == 8 ? b8 : b8 >> 1;
}

// Plane xx of f, clamped into the grid: a uniform pointer. Values read from a
// clamped plane only feed terms that the boundary masks drop.
template <typename T>
__device__ __forceinline__ T* plane_at(T* f, int xx, const Grid& g) {
    return f + (int64_t)min(max(xx, 0), g.nx - 1) * g.plane();
}

__device__ inline CgTile cg_tile(const Grid& g, int x0, int x1, int y0, int z0) {
    CgTile t;
    const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tz;
    t.x0 = x0;
    t.x1 = x1;
    t.y = y0 + ty;
    t.z = z0 + tz;
    t.inyz = t.y < g.ny && t.z < g.nz;
    t.yz = t.inyz ? (uint32_t)(t.y * g.nz + t.z) : 0u;
    t.yzb = t.yz * (uint32_t)sizeof(double);
    // halo: rows y0-2, y0-1, y0+BY, y0+BY+1 (4 x 64 points), columns z0-2, z0-1, z0+64, z0+65 (4 x BY)
    int hy = -1, hz = -1;
    t.ly = t.lz = 0;
    if (tid < 256) {
        const int rr = tid >> 6;
        hy = rr < 2 ? y0 - 2 + rr : y0 + kCgBY + rr - 2;
        t.ly = rr < 2 ? rr : kCgBY + rr;
        hz = z0 + (tid & 63);
        t.lz = (tid & 63) + 2;
    } else if (tid < 256 + 4 * kCgBY) {
        const int k = tid - 256, cc = k / kCgBY, yy = k % kCgBY;
        hy = y0 + yy;
        t.ly = yy + 2;
        hz = cc < 2 ? z0 - 2 + cc : z0 + 64 + cc - 2;
        t.lz = cc < 2 ? cc : 64 + cc;
    }
    t.halo_thread = tid < 256 + 4 * kCgBY;
    t.halo_in = t.halo_thread && hy >= 0 && hy < g.ny && hz >= 0 && hz < g.nz;
    t.hyzb = t.halo_in ? (uint32_t)(hy * g.nz + hz) * (uint32_t)sizeof(double) : 0u;
    return t;
}

// A p at an interior voxel: x terms from the queue (a = x-2, c = x, e = x+2),
// y and z terms from the LDS plane; rows next to the wall drop G on the
// boundary layer (solver.py:1024-1026 with gradient_center zero there).
template <typename V>
__device__ inline double cg_apply(const Grid& g, int x, int y, int z, double a, double c, double e,
                                  const V (*L)[kCgLZ], int r, int q, double qx, double qy, double qz) {
    // the four LDS reads are unconditional (always inside the tile's halo), so
    // they issue together and the masks below are selects, not branches
    const double yp = (double)L[r + 2][q], ym = (double)L[r - 2][q], zp = (double)L[r][q + 2],
                 zm = (double)L[r][q - 2];
    const double tx = ((x + 1 < g.nx - 1) ? (e - c) : 0.0) - ((x - 1 > 0) ? (c - a) : 0.0);
    const double ty = ((y + 1 < g.ny - 1) ? (yp - c) : 0.0) - ((y - 1 > 0) ? (c - ym) : 0.0);
    const double tz = ((z + 1 < g.nz - 1) ? (zp - c) : 0.0) - ((z - 1 > 0) ? (c - zm) : 0.0);
    return -(tx * qx + ty * qy + tz * qz);
}

// Prefetch distances of the CG tiles, in x planes: a plane's loads are issued
// this many iterations before the one that consumes them, and nothing is
// computed on a loaded value before then, so no wait for memory falls inside
// the iteration that issued it. One 16-wave block holds a CU, so deeper rings
// keep more loads in flight. The dir tile's rings hold eight planes (1..7),
// the step tile's p ring eight planes from x-2 (1..3).
#ifndef SW_CG_DIST_DIR
#define SW_CG_DIST_DIR 3
#endif
#ifndef SW_CG_DIST
#define SW_CG_DIST 3
#endif
constexpr int kCgDistDir = SW_CG_DIST_DIR;
constexpr int kCgDist = SW_CG_DIST;
static_assert(kCgDistDir >= 1 && kCgDistDir <= 7, "CG dir prefetch distance");
static_assert(kCgDist >= 1 && kCgDist <= 3, "CG step prefetch distance");

template <int K>
using Ic = std::integral_constant<int, K>;

// The x loop of the CG tiles is unrolled by kCgU planes so that every ring
// slot is a compile-time index (the rings live in registers and rotate
// without moves). Loads issued near the end of one unrolled trip are
// consumed early in the next, and the compiler waits for every load in
// flight at the loop header (vmcnt(0)): a longer trip drains the queues less
// often.
#ifndef SW_CG_UNROLL
#define SW_CG_UNROLL 16
#endif
constexpr int kCgU = SW_CG_UNROLL, kCgM = kCgU - 1;
static_assert(kCgU == 8 || kCgU == 16, "CG unroll");

template <typename F, int... Ks>
__device__ __forceinline__ bool march_trip(int x, int x1, F& body, std::integer_sequence<int, Ks...>) {
    // left to right, stopping at the first plane past the chunk
    return ((x + Ks < x1 && (body(Ic<Ks>{}, x + Ks), true)) && ...);
}

// Sub-chunks of kCgSub x planes: a tile hands its dot-product share to its
// sink at the end of each one (and at x1), so partial sums can cover the same
// planes whatever the length of the work item that marched over them.
constexpr int kCgSub = 48;
static_assert(kCgSub % kCgU == 0, "CG sub-chunk");

// A sink that keeps one running share for the whole tile.
struct CgNoSink {
    __device__ void operator()(double&, int) const {}
};

// body(Ic<k>{}, x) for x = x0 .. x1-1 with k = (x - x0) mod kCgU; sink(j) after
// the last plane of sub-chunk j (planes x0 + j kCgSub ..), uniform over the block.
template <typename F, typename G>
__device__ __forceinline__ void march_cg(int x0, int x1, F&& body, G&& sink) {
    for (int x = x0; x < x1; x += kCgU) {
        const bool full = march_trip(x, x1, body, std::make_integer_sequence<int, kCgU>{});
        if (x + kCgU >= x1 || (x + kCgU - x0) % kCgSub == 0) sink((x - x0) / kCgSub);
        if (!full) break;
    }
}

// r and p_old of one point, as loaded.
template <typename V>
struct CgRaw {
    V r, q;
};

// p = r + beta p_old (one fma, rounded to V, so every recomputation is
// bit-identical to the stored p), A p. Writes p over the tile; returns this
// thread's share of p.Ap (what the sink left of it: sink(share, j) may take
// and clear the share at the end of sub-chunk j).
template <typename V, typename Sink = CgNoSink>
__device__ inline double cg_dir_tile(const V* __restrict__ r, const V* __restrict__ p_old, V* __restrict__ p,
                                     const Grid& g, double beta, double qx, double qy, double qz, const CgTile& t,
                                     CgLdsT<V>& lds, Sink sink = Sink{}) {
    const int r0 = threadIdx.y + 2, q0 = threadIdx.x + 2;
    const uint32_t cb = vb<V>(t.yzb), hb = vb<V>(t.hyzb);
    // plain loads: these lines also serve the halos of the neighbouring tiles
    auto raw = [&](int xx, uint32_t o) {
        return CgRaw<V>{at_byte(plane_at(r, xx, g), o), at_byte(plane_at(p_old, xx, g), o)};
    };
    auto center = [&](int xx) { return raw(xx, cb); };
    // halo loads are unconditional (lanes without a halo point read offset 0),
    // so no branch separates a load from the wait that precedes its use
    auto halo = [&](int xx) { return raw(xx, hb); };
    auto pv = [&](const CgRaw<V>& v) { return as_stored<V>(fma(beta, (double)v.q, (double)v.r)); };
    const int x0 = t.x0, x1 = t.x1;
    constexpr int D = kCgDistDir;
    double pq[kCgU];    // p of this column at plane x0 + k in slot k & kCgM (planes x-2 .. x+2 live)
    CgRaw<V> rq[kCgU];  // raw loads of this column in flight, slot k & kCgM (planes x+3 .. x+2+D)
    CgRaw<V> hq[kCgU];  // raw loads of this thread's halo point, slot k & kCgM (planes x .. x+D)
#pragma unroll
    for (int k = -2; k < 2; ++k) pq[k & kCgM] = pv(center(x0 + k));
#pragma unroll
    for (int k = 2; k < 2 + D; ++k) rq[k & kCgM] = center(x0 + k);
#pragma unroll
    for (int k = 0; k < D; ++k) hq[k & kCgM] = halo(x0 + k);
    // plane x0 is staged ahead of the loop; each iteration stages the next
    // plane right after its barrier, so the LDS stores overlap this plane's work
    lds[0][r0][q0] = (V)pq[0];
    if (t.halo_thread) lds[0][t.ly][t.lz] = (V)pv(hq[0]);
    double s = 0.0;
    march_cg(x0, x1, [&](auto kc, int x) {
        constexpr int K = decltype(kc)::value;
        constexpr int B = K & 1;
        const double e = pv(rq[(K + 2) & kCgM]);  // plane x+2, loaded D iterations ago
        pq[(K + 2) & kCgM] = e;
        const double a = pq[(K + kCgU - 2) & kCgM], c = pq[K & kCgM];
        rq[(K + 2 + D) & kCgM] = center(x + 2 + D);
        hq[(K + D) & kCgM] = halo(x + D);
        __syncthreads();
        // plane x+1 into the other buffer: its last reads (plane x-1) ended before the barrier
        lds[B ^ 1][r0][q0] = (V)pq[(K + 1) & kCgM];
        if (t.halo_thread) lds[B ^ 1][t.ly][t.lz] = (V)pv(hq[(K + 1) & kCgM]);
        if (!t.inyz) return;
        const bool in = g.interior(x, t.y, t.z);
        if (in) s += c * cg_apply<V>(g, x, t.y, t.z, a, c, e, lds[B], r0, q0, qx, qy, qz);
        // float64 p goes out non-temporal (209 -> 202 us per iteration at 256^3);
        // float32 p keeps plain stores (138 us, against 144 non-temporal)
        if constexpr (sizeof(V) == 8)
            store_out(&at_byte(plane_at(p, x, g), cb), in ? (V)c : (V)0);
        else
            at_byte(plane_at(p, x, g), cb) = in ? (V)c : (V)0;
    }, [&](int j) { sink(s, j); });
    return s;
}

// phi updates of the step tile. PHI 1: phi += alpha p every iteration.
// Deferred (PHI 0, then PHI 2): an odd iteration k leaves phi alone, and
// iteration k+1 adds both of its terms, phi = (phi + alpha_k p_k) +
// alpha_{k+1} p_{k+1}, as two fmas in that order: bit-identical to updating
// every iteration, with phi read and written every other iteration only (p_k
// is still in the other direction buffer). A solve that stops after an odd
// iteration adds its pending term with cg_flush_phi.
enum : int { kPhiSkip = 0, kPhiEach = 1, kPhiPair = 2 };

// phi update(s) per PHI, r -= alpha A p (A p recomputed from p) over the
// tile; returns this thread's share of r.r (over r as stored; the sink works
// as in cg_dir_tile). r and phi are
// read and written once here (non-temporal, except float32 r); p keeps plain loads, its lines
// serve the halos of the neighbouring tiles. Loaded values stay in V until
// the iteration that consumes them.
template <typename V, int PHI, typename Sink = CgNoSink>
__device__ inline double cg_step_tile(double* __restrict__ phi, V* __restrict__ r, const V* __restrict__ p,
                                      const V* __restrict__ p_prev, const Grid& g, double alpha, double alpha_prev,
                                      double qx, double qy, double qz, const CgTile& t, CgLdsT<V>& lds,
                                      Sink sink = Sink{}) {
    const int r0 = threadIdx.y + 2, q0 = threadIdx.x + 2;
    const uint32_t cb = vb<V>(t.yzb), hb = vb<V>(t.hyzb);
    auto center = [&](int xx) { return at_byte(plane_at(p, xx, g), cb); };
    auto halo = [&](int xx) { return at_byte(plane_at(p, xx, g), hb); };
    // float32 r is loaded and stored with plain accesses: measured at 256^3,
    // 132.7 us per iteration against 138.0 with non-temporal ones. float64 r
    // keeps non-temporal accesses; phi is faster non-temporal in both forms
    // (plain phi stores: 141 against 136 us, 206 against 202 us in float64)
    auto once_r = [&](int xx) -> V {
        if constexpr (sizeof(V) == 8)
            return load_once(&at_byte(plane_at(r, xx, g), cb));
        else
            return at_byte(plane_at(r, xx, g), cb);
    };
    auto once_phi = [&](int xx) { return load_once(&at_byte(plane_at(phi, xx, g), t.yzb)); };
    auto once_pp = [&](int xx) { return load_once(&at_byte(plane_at(p_prev, xx, g), cb)); };
    const int x0 = t.x0, x1 = t.x1;
    V pq[kCgU];    // p of this column, slot k & kCgM (planes x-2 .. x+2+D live)
    V hq[4];       // p at this thread's halo point, slot k & 3 (planes x .. x+D)
    V rq[4];       // r of this column, slot k & 3
    double fq[4];  // phi of this column, slot k & 3 (PHI != 0)
    V oq[4];       // p_prev of this column, slot k & 3 (PHI == 2)
#pragma unroll
    for (int k = -2; k < 2 + kCgDist; ++k) pq[k & kCgM] = center(x0 + k);
#pragma unroll
    for (int k = 0; k < kCgDist; ++k) {
        hq[k & 3] = halo(x0 + k);
        rq[k & 3] = once_r(x0 + k);
        if constexpr (PHI != kPhiSkip) fq[k & 3] = once_phi(x0 + k);
        if constexpr (PHI == kPhiPair) oq[k & 3] = once_pp(x0 + k);
    }
    // plane x0 is staged ahead of the loop; each iteration stages the next
    // plane right after its barrier, so the LDS stores overlap this plane's work
    lds[0][r0][q0] = pq[0];
    if (t.halo_thread) lds[0][t.ly][t.lz] = hq[0];
    double s = 0.0;
    march_cg(x0, x1, [&](auto kc, int x) {
        constexpr int K = decltype(kc)::value;
        constexpr int B = K & 1;
        const V cv = pq[K & kCgM];
        const double a = (double)pq[(K + kCgU - 2) & kCgM], c = (double)cv, e = (double)pq[(K + 2) & kCgM];
        pq[(K + 2 + kCgDist) & kCgM] = center(x + 2 + kCgDist);
        hq[(K + kCgDist) & 3] = halo(x + kCgDist);
        rq[(K + kCgDist) & 3] = once_r(x + kCgDist);
        if constexpr (PHI != kPhiSkip) fq[(K + kCgDist) & 3] = once_phi(x + kCgDist);
        if constexpr (PHI == kPhiPair) oq[(K + kCgDist) & 3] = once_pp(x + kCgDist);
        __syncthreads();
        // plane x+1 into the other buffer: its last reads (plane x-1) ended before the barrier
        lds[B ^ 1][r0][q0] = pq[(K + 1) & kCgM];
        if (t.halo_thread) lds[B ^ 1][t.ly][t.lz] = hq[(K + 1) & 3];
        if (!t.inyz || !g.interior(x, t.y, t.z)) return;
        const double ap = cg_apply<V>(g, x, t.y, t.z, a, c, e, lds[B], r0, q0, qx, qy, qz);
        if constexpr (PHI == kPhiEach)
            store_out(&at_byte(plane_at(phi, x, g), t.yzb), fma(alpha, c, fq[K & 3]));
        if constexpr (PHI == kPhiPair)
            store_out(&at_byte(plane_at(phi, x, g), t.yzb),
                      fma(alpha, c, fma(alpha_prev, (double)oq[K & 3], fq[K & 3])));
        const V rv = (V)((double)rq[K & 3] - alpha * ap);
        if constexpr (sizeof(V) == 8)
            store_out(&at_byte(plane_at(r, x, g), cb), rv);
        else
            at_byte(plane_at(r, x, g), cb) = rv;
        s += (double)rv * (double)rv;
    }, [&](int j) { sink(s, j); });
    return s;
}

// The pending phi term of a solve whose last iteration k was odd: phi +=
// alpha_k p_k over a tile (p is zero on the boundary layer, so boundary
// voxels keep phi = 0).
template <typename V>
__device__ inline void cg_flush_phi_tile(double* __restrict__ phi, const V* __restrict__ p, const Grid& g,
                                         double alpha, const CgTile& t) {
    if (!t.inyz) return;
    for (int x = t.x0; x < t.x1; ++x) {
        if (!g.interior(x, t.y, t.z)) continue;
        auto& f = at_byte(plane_at(phi, x, g), t.yzb);
        f = fma(alpha, (double)at_byte(plane_at(p, x, g), vb<V>(t.yzb)), f);
    }
}

// r0 = -div_c(mask f) over a tile (no LDS: each f component is read at +-1
// along its own axis only), p_old = 0, phi = 0. Returns r.r and max|div|.
template <typename T, typename V>
__device__ inline void cg_init_tile(const T* __restrict__ fx, const T* __restrict__ fy, const T* __restrict__ fz,
                                    double* __restrict__ phi, V* __restrict__ r, V* __restrict__ p_old,
                                    const Grid& g, double i2hx, double i2hy, double i2hz, const CgTile& t, double& s,
                                    double& mx) {
    if (!t.inyz) return;
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    for (int x = t.x0; x < t.x1; ++x) {
        const int64_t i = (int64_t)x * P + t.yz;
        double rv = 0.0;
        if (g.interior(x, t.y, t.z)) {
            // neighbours on the boundary layer are masked to zero (solver.py:666-668)
            const double xp = (x + 1 < g.nx - 1) ? (double)fx[i + P] : 0.0;
            const double xm = (x - 1 > 0) ? (double)fx[i - P] : 0.0;
            const double yp = (t.y + 1 < g.ny - 1) ? (double)fy[i + R] : 0.0;
            const double ym = (t.y - 1 > 0) ? (double)fy[i - R] : 0.0;
            const double zp = (t.z + 1 < g.nz - 1) ? (double)fz[i + 1] : 0.0;
            const double zm = (t.z - 1 > 0) ? (double)fz[i - 1] : 0.0;
            const double div = (xp - xm) * i2hx + (yp - ym) * i2hy + (zp - zm) * i2hz;
            rv = as_stored<V>(-div);
            s += rv * rv;
            mx = fmax(mx, fabs(div));
        }
        r[i] = (V)rv;
        p_old[i] = (V)0;
        phi[i] = 0.0;
    }
}

// f_S = mask (mask f - grad_c phi) over a tile; optionally out = scale f_S inv_rho.
template <typename T>
__device__ inline void cg_grad_sub_tile(T* __restrict__ fx, T* __restrict__ fy, T* __restrict__ fz,
                                        const double* __restrict__ phi, const Grid& g, double i2hx, double i2hy,
                                        double i2hz, T* __restrict__ ox, T* __restrict__ oy, T* __restrict__ oz,
                                        const T* __restrict__ inv_rho, double scale, const CgTile& t) {
    if (!t.inyz) return;
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    for (int x = t.x0; x < t.x1; ++x) {
        const int64_t i = (int64_t)x * P + t.yz;
        double a = 0.0, b = 0.0, c = 0.0;
        if (g.interior(x, t.y, t.z)) {
            a = (double)fx[i] - (phi[i + P] - phi[i - P]) * i2hx;
            b = (double)fy[i] - (phi[i + R] - phi[i - R]) * i2hy;
            c = (double)fz[i] - (phi[i + 1] - phi[i - 1]) * i2hz;
        }
        fx[i] = (T)a;
        fy[i] = (T)b;
        fz[i] = (T)c;
        if (ox) {
            const double w = scale * (inv_rho ? (double)inv_rho[i] : 1.0);
            ox[i] = (T)(a * w);
            oy[i] = (T)(b * w);
            oz[i] = (T)(c * w);
        }
    }
}

}  // namespace sw
