// Static elasticity solve on gfx950: elastic_solve_static (solver.py:569-625).
//
// Conjugate gradients on (-mu lap) u = bz with a zero Dirichlet boundary, in
// float32 as the reference runs it: float32 vectors and scalars, the masked
// updates in the reference's order, and every dot product summed in NumPy's
// float32 order, so each iterate, each scalar and the iteration count are
// bit-identical to the reference's.
//
// NumPy's np.sum of a C-contiguous float32 array takes the array in buffers
// of 8192 elements and adds each buffer's pairwise sum to a running total
// (from 0) in order. The pairwise sum of m elements splits at
// n2 = m/2 - (m/2) % 8 until m <= 128; a leaf of at least 8 elements runs
// eight interleaved accumulators, combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)),
// then adds its leftover elements in order; a shorter one is a plain loop.
//
// Layout: one 512-thread block per 8192-element chunk of the flat (C-order)
// arrays, chunks of one XCD contiguous so the x-neighbour planes a chunk reads
// sit in that XCD's L2. Each thread computes 16 voxels (coalesced, stride
// 512) and writes their products into LDS; eight-lane groups then sum the
// leaves of the chunk's tree (the eight accumulators are the eight lanes, the
// accumulator tree is a xor-1/2/4 butterfly) and the internal nodes are added
// level by level from a plan built on the host. Each chunk's sum goes to a
// slot; the last block adds the slots in chunk order. One iteration is two
// kernels:
//   dir:  p = (r + beta p_old) mask at the 7 points of A, stores p, p . Ap
//   step: Ap recomputed from p (same operations), u += alpha p, r -= alpha Ap
//         (both masked), r . r, and the stopping rules.
#include <cmath>
#include <cstring>
#include <vector>

#include "march_cg.h"
#include "sw_common.h"

#pragma clang fp contract(off)

namespace sw {
namespace {

constexpr int kChunk = 8192;  // NumPy's reduction buffer (np.getbufsize())
constexpr int kThreads = 512;
constexpr int kPerThread = kChunk / kThreads;
constexpr int kGroups = kThreads / 8;  // eight-lane groups, one leaf at a time
constexpr int kMaxLeaves = 96;         // a chunk's tree has at most 65 leaves
constexpr int kMaxLevels = 8;          // and at most 7 levels of internal nodes
constexpr int kPadded = kChunk + kChunk / 32;
constexpr float kEps32 = 1.1920928955078125e-07f;  // np.finfo(np.float32).eps

// Voxels of a thread whose loads are issued together, and an optional cap on
// VGPRs (waves per SIMD) so more blocks fit a CU.
#ifndef SW_STATIC_UNROLL
#define SW_STATIC_UNROLL 4
#endif
constexpr int kUnroll = SW_STATIC_UNROLL;
#ifdef SW_STATIC_WAVES
#define SW_STATIC_OCC __attribute__((amdgpu_waves_per_eu(SW_STATIC_WAVES)))
#else
#define SW_STATIC_OCC
#endif

// The pairwise tree of one chunk length: leaves in order, then internal nodes
// numbered n_leaf, n_leaf + 1, ... level by level (children in earlier levels).
struct SumPlan {
    int n_leaf, n_level;
    int16_t level_end[kMaxLevels];  // internal nodes [level_end[l-1], level_end[l]) form level l
    int16_t leaf_start[kMaxLeaves], leaf_len[kMaxLeaves];
    int16_t node_l[kMaxLeaves], node_r[kMaxLeaves];
};

struct StaticState {
    float rs_old, rs0, alpha, beta, rel, tol_f;
    double tol_d;
    long long iters, max_iters;
    int stop, reason, tol_is_f32, pad0;
    unsigned int counter;  // slots published in the current kernel
    int pad;
};

// Coordinate step of a flat index advanced by kThreads.
struct Walk {
    int dx, dy, dz;
};

__device__ __forceinline__ int lds_idx(int e) { return e + (e >> 5); }

__device__ __forceinline__ void coords(int64_t i, const Grid& g, int& x, int& y, int& z) {
    const int64_t P = g.plane();
    x = (int)(i / P);
    const int64_t rem = i - (int64_t)x * P;
    y = (int)(rem / g.nz);
    z = (int)(rem - (int64_t)y * g.nz);
}

__device__ __forceinline__ void advance(int& x, int& y, int& z, const Grid& g, const Walk& w) {
    z += w.dz;
    const int carry = z >= g.nz;
    if (carry) z -= g.nz;
    y += w.dy + carry;
    if (y >= g.ny) {
        y -= g.ny;
        x += 1;
    }
    x += w.dx;
}

// One voxel of a thread's walk. Its loads are unconditional, so the loads of
// several voxels issue together: a voxel past the chunk's end reads the
// chunk's first voxel (and stores nothing), and a voxel on the boundary layer
// reads itself in place of its neighbours (its Laplacian is 0, selected after).
struct Vox {
    int e;          // element in the chunk
    bool live, in;  // inside the chunk; on the interior
    int64_t i;      // flat index
    int64_t oP, oR, o1;
};

__device__ __forceinline__ Vox vox_at(int e, int m, int64_t base, int x, int y, int z, const Grid& g) {
    Vox v;
    v.e = e;
    v.live = e < m;
    v.i = base + (v.live ? e : 0);
    v.in = v.live && g.interior(x, y, z);
    v.oP = v.in ? g.plane() : 0;
    v.oR = v.in ? (int64_t)g.nz : 0;
    v.o1 = v.in ? 1 : 0;
    return v;
}

// Chunk of work item vb (0 <= vb < nchunks): the items of one XCD (vb % 8,
// the XCD of every block that runs them when the grid is a multiple of 8)
// cover a contiguous range of chunks.
__device__ __forceinline__ int chunk_of(int vb, int nchunks) {
    const int q = nchunks >> 3, rem = nchunks & 7, xcd = vb & 7;
    return xcd * q + min(xcd, rem) + (vb >> 3);
}

// ((c+ - 2c) + c-) / h^2 per axis, summed x, y, z (laplacian_center, solver.py:686-695).
// v(offset, q) is the neighbour at that flat offset, q = 0..5 for x+, x-, y+, y-, z+, z-.
// On the boundary layer the offsets are 0 and v returns the centre value c at
// every tap, so each axis term is ((c - 2c) + c) / h^2 = (-c + c) / h^2 = +0
// exactly (finite c): the reference's zero Laplacian there, with no branch
// that would serialise the loads of successive voxels.
template <typename V>
__device__ __forceinline__ float lap7(V v, float c, const Vox& t, float hx2, float hy2, float hz2) {
    const float c2 = 2.0f * c;
    const float tx = ((v(t.oP, 0) - c2) + v(-t.oP, 1)) / hx2;
    const float ty = ((v(t.oR, 2) - c2) + v(-t.oR, 3)) / hy2;
    const float tz = ((v(t.o1, 4) - c2) + v(-t.o1, 5)) / hz2;
    return (tx + ty) + tz;
}

// The chunk's sum in NumPy's pairwise order from the products in LDS
// (prod[lds_idx(e)], e < the plan's length). Valid in every thread.
__device__ float plan_sum(const float* prod, float* val, const SumPlan& pl) {
    const int tid = threadIdx.x, grp = tid >> 3, k = tid & 7;
    const int n_leaf = pl.n_leaf;
    for (int j = grp; j < n_leaf; j += kGroups) {
        const int s = pl.leaf_start[j], len = pl.leaf_len[j];
        float res;
        if (len < 8) {
            res = 0.0f;
            for (int q = 0; q < len; ++q) res = res + prod[lds_idx(s + q)];
        } else {
            const int full = len & ~7;
            float a = prod[lds_idx(s + k)];
            for (int o = 8; o < full; o += 8) a = a + prod[lds_idx(s + o + k)];
            a = a + __shfl_xor(a, 1, 8);  // r0+r1, r2+r3, ...
            a = a + __shfl_xor(a, 2, 8);  // (r0+r1)+(r2+r3), ...
            a = a + __shfl_xor(a, 4, 8);
            res = a;
            for (int q = full; q < len; ++q) res = res + prod[lds_idx(s + q)];
        }
        if (k == 0) val[j] = res;
    }
    __syncthreads();
    int lo = 0;
    for (int l = 0; l < pl.n_level; ++l) {
        const int hi = pl.level_end[l];
        for (int q = lo + tid; q < hi; q += kThreads) val[n_leaf + q] = val[pl.node_l[q]] + val[pl.node_r[q]];
        lo = hi;
        __syncthreads();
    }
    return val[n_leaf > 1 ? n_leaf + lo - 1 : 0];
}

// In the last block: 0 + slot[0] + slot[1] + ... in chunk order (valid in thread 0).
__device__ float ordered_total(const double* slots, int n, float* buf) {
    float tot = 0.0f;
    for (int b0 = 0; b0 < n; b0 += kChunk) {
        const int cnt = min(kChunk, n - b0);
        for (int q = threadIdx.x; q < cnt; q += kThreads) buf[q] = (float)load_coherent(slots + b0 + q);
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll 32
            for (int q = 0; q < cnt; ++q) tot = tot + buf[q];
        }
        __syncthreads();
    }
    return tot;
}

// The block's chunks (work items blockIdx.x, + gridDim.x, ...): body(base, m)
// fills prod with the products of the m voxels from flat index base; each
// chunk's pairwise sum goes to its slot. Returns true in the block that
// finishes last, with the ordered total of all slots in thread 0: one release
// fence and one atomic per block, not per chunk.
template <typename F>
__device__ bool chunk_sums(int nchunks, const Grid& g, const SumPlan* __restrict__ plans, double* slots,
                           StaticState* st, float* prod, float* val, float* total, F&& body) {
    for (int vb = blockIdx.x; vb < nchunks; vb += gridDim.x) {
        const int c = chunk_of(vb, nchunks);
        const int64_t base = (int64_t)c * kChunk;
        const int m = (int)min<int64_t>(kChunk, g.size() - base);
        body(base, m);
        __syncthreads();
        const float s = plan_sum(prod, val, plans[m == kChunk ? 0 : 1]);
        if (threadIdx.x == 0) slots[c] = (double)s;
        __syncthreads();  // prod and val are reused by the next chunk
    }
    __shared__ int last;
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&st->counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    *total = ordered_total(slots, nchunks, prod);
    return true;
}

__global__ void static_reset_kernel(StaticState* st, double tol, int tol_is_f32, long long max_iters) {
    st->rs_old = st->rs0 = st->alpha = st->beta = 0.0f;
    st->rel = __int_as_float(0x7fc00000);  // no iteration yet
    st->tol_d = tol;
    st->tol_f = (float)tol;
    st->tol_is_f32 = tol_is_f32;
    st->iters = 0;
    st->max_iters = max_iters;
    st->stop = max_iters <= 0 ? 1 : 0;
    st->reason = SW_CG_MAX_ITERS;
    st->counter = 0u;
}

template <bool FIELD>
__device__ __forceinline__ float mu_at(const float* __restrict__ mu, float mu_s, int64_t i) {
    return FIELD ? mu[i] : mu_s;
}

// The voxels of one thread: e = threadIdx.x + k * kThreads, k < kPerThread.
// FULL chunks need no liveness test, so nothing but the walk separates the
// voxels and the unrolled loop issues the loads of several voxels together.
template <bool FULL, typename F>
__device__ __forceinline__ void for_voxels(int m, int64_t base, const Grid& g, const Walk& w, F&& body) {
    int x, y, z;
    coords(base + threadIdx.x, g, x, y, z);
#pragma unroll kUnroll
    for (int k = 0; k < kPerThread; ++k) {
        Vox t = vox_at(threadIdx.x + k * kThreads, FULL ? kChunk : m, base, x, y, z, g);
        if (FULL) t.live = true;
        body(t, x, y, z);
        advance(x, y, z, g, w);
    }
}

// Four consecutive voxels of a z row per thread and pass (grids with nz % 4
// == 0 and 16-byte aligned arrays, whole chunks): the centre, y and x taps are
// float4 loads and the z taps of the inner voxels are the neighbouring centre
// values, so a voxel costs about a quarter of the load instructions. The
// arithmetic per voxel is the same as above.
constexpr int kQuad = 4 * kThreads;  // voxels per pass

struct F4 {
    float v[4];
};

__device__ __forceinline__ F4 ld4(const float* __restrict__ p, int64_t i) {
    const float4 t = *reinterpret_cast<const float4*>(p + i);
    return F4{{t.x, t.y, t.z, t.w}};
}

__device__ __forceinline__ void st4(float* __restrict__ p, int64_t i, const F4& a) {
    *reinterpret_cast<float4*>(p + i) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
}

struct Quad {
    int e;           // first element in the chunk
    int64_t i;       // its flat index
    int z0;          // its z (a multiple of 4)
    bool in[4];      // interior voxels
    int64_t oP, oR;  // x / y neighbour offsets: 0 unless one of the four is interior
    int64_t ozm, ozp;  // z-1 of voxel 0 and z+1 of voxel 3 (0 unless that voxel is interior)
};

__device__ __forceinline__ Quad quad_at(int e, int64_t base, int x, int y, int z0, const Grid& g) {
    Quad q;
    q.e = e;
    q.i = base + e;
    q.z0 = z0;
    const bool xy = x > 0 && y > 0 && x < g.nx - 1 && y < g.ny - 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) q.in[k] = xy && z0 + k > 0 && z0 + k < g.nz - 1;
    const bool any = q.in[0] || q.in[1] || q.in[2] || q.in[3];
    q.oP = any ? g.plane() : 0;
    q.oR = any ? (int64_t)g.nz : 0;
    q.ozm = q.in[0] ? -1 : 0;  // voxel 0 interior: z0 >= 4, the row continues below
    q.ozp = q.in[3] ? 4 : 0;   // voxel 3 interior: z0 + 4 < nz, the row continues above
    return q;
}

// lap7 of one voxel from its taps; a boundary voxel's taps become the voxel
// itself, which gives +0 exactly as in lap7.
__device__ __forceinline__ float lap_taps(float c, float xp, float xm, float yp, float ym, float zp, float zm, bool in,
                                          float hx2, float hy2, float hz2) {
    if (!in) xp = xm = yp = ym = zp = zm = c;
    const float c2 = 2.0f * c;
    const float tx = ((xp - c2) + xm) / hx2;
    const float ty = ((yp - c2) + ym) / hy2;
    const float tz = ((zp - c2) + zm) / hz2;
    return (tx + ty) + tz;
}

template <typename F>
__device__ __forceinline__ void for_quads(int64_t base, const Grid& g, const Walk& w4, F&& body) {
    int x, y, z;
    coords(base + 4 * threadIdx.x, g, x, y, z);
#pragma unroll
    for (int k = 0; k < kChunk / kQuad; ++k) {
        body(quad_at(4 * threadIdx.x + k * kQuad, base, x, y, z, g), x, y);
        advance(x, y, z, g, w4);
    }
}

template <bool FIELD>
__device__ __forceinline__ F4 mu4(const float* __restrict__ mu, float mu_s, int64_t i) {
    if (FIELD) return ld4(mu, i);
    return F4{{mu_s, mu_s, mu_s, mu_s}};
}

template <bool FIELD>
__device__ __forceinline__ void init_quads(const float* __restrict__ rhs, const float* __restrict__ mu, float mu_s,
                                           float* __restrict__ u, float* __restrict__ r, const Grid& g,
                                           const Walk& w4, int64_t base, float* prod) {
    for_quads(base, g, w4, [&](const Quad& q, int, int) {
        const F4 b = ld4(rhs, q.i), mv = mu4<FIELD>(mu, mu_s, q.i);
        F4 rn;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rn.v[k] = (b.v[k] - (-mv.v[k] * 0.0f)) * (q.in[k] ? 1.0f : 0.0f);
            prod[lds_idx(q.e + k)] = rn.v[k] * rn.v[k];
        }
        st4(r, q.i, rn);
        st4(u, q.i, F4{{0.0f, 0.0f, 0.0f, 0.0f}});
    });
}

template <bool FIRST, bool FIELD>
__device__ __forceinline__ void dir_quads(const float* __restrict__ r, const float* __restrict__ p_old,
                                          float* __restrict__ p, const float* __restrict__ mu, float mu_s,
                                          const Grid& g, const Walk& w4, float beta, float hx2, float hy2, float hz2,
                                          int64_t base, float* prod) {
    for_quads(base, g, w4, [&](const Quad& q, int x, int y) {
        // r + beta p_old (unmasked) at four voxels from offset off, and at one voxel
        auto pv = [&](int64_t off) {
            F4 o = ld4(r, q.i + off);
            if (!FIRST) {
                const F4 qq = ld4(p_old, q.i + off);
#pragma unroll
                for (int k = 0; k < 4; ++k) o.v[k] = o.v[k] + beta * qq.v[k];
            }
            return o;
        };
        auto ps = [&](int64_t j) { return FIRST ? r[j] : r[j] + beta * p_old[j]; };
        // the first iteration copies p = r unmasked; later ones mask every value
        auto msk = [&](bool b) { return (FIRST || b) ? 1.0f : 0.0f; };
        F4 C = pv(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) C.v[k] = C.v[k] * msk(q.in[k]);
        const F4 XP = pv(q.oP), XM = pv(-q.oP), YP = pv(q.oR), YM = pv(-q.oR);
        const float ZM = ps(q.i + q.ozm), ZP = ps(q.i + q.ozp);
        // taps of an interior voxel: interior unless on the wall of their axis
        const float mxp = msk(x + 1 < g.nx - 1), mxm = msk(x > 1), myp = msk(y + 1 < g.ny - 1), mym = msk(y > 1);
        const float mzp = msk(q.z0 + 4 < g.nz - 1), mzm = msk(q.z0 > 1);
        const F4 mv = mu4<FIELD>(mu, mu_s, q.i);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // the inner z taps are the neighbouring centre values, already masked as theirs
            const float zp = k < 3 ? C.v[k + 1] : ZP * mzp;
            const float zm = k > 0 ? C.v[k - 1] : ZM * mzm;
            const float lap = lap_taps(C.v[k], XP.v[k] * mxp, XM.v[k] * mxm, YP.v[k] * myp, YM.v[k] * mym, zp, zm,
                                       q.in[k], hx2, hy2, hz2);
            prod[lds_idx(q.e + k)] = C.v[k] * (-mv.v[k] * lap);
        }
        st4(p, q.i, C);
    });
}

template <bool FIELD>
__device__ __forceinline__ void step_quads(float* __restrict__ u, float* __restrict__ r, const float* __restrict__ p,
                                           const float* __restrict__ mu, float mu_s, const Grid& g, const Walk& w4,
                                           float alpha, float hx2, float hy2, float hz2, int64_t base, float* prod) {
    for_quads(base, g, w4, [&](const Quad& q, int, int) {
        const F4 C = ld4(p, q.i), XP = ld4(p, q.i + q.oP), XM = ld4(p, q.i - q.oP);
        const F4 YP = ld4(p, q.i + q.oR), YM = ld4(p, q.i - q.oR);
        const float ZM = p[q.i + q.ozm], ZP = p[q.i + q.ozp];
        const F4 U = ld4(u, q.i), R = ld4(r, q.i), mv = mu4<FIELD>(mu, mu_s, q.i);
        F4 un, rn;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float zp = k < 3 ? C.v[k + 1] : ZP;
            const float zm = k > 0 ? C.v[k - 1] : ZM;
            const float lap =
                lap_taps(C.v[k], XP.v[k], XM.v[k], YP.v[k], YM.v[k], zp, zm, q.in[k], hx2, hy2, hz2);
            const float ap = -mv.v[k] * lap;
            const float mk = q.in[k] ? 1.0f : 0.0f;
            un.v[k] = (U.v[k] + alpha * C.v[k]) * mk;
            rn.v[k] = (R.v[k] - alpha * ap) * mk;
            prod[lds_idx(q.e + k)] = rn.v[k] * rn.v[k];
        }
        st4(u, q.i, un);
        st4(r, q.i, rn);
    });
}

template <bool FIELD, bool FULL>
__device__ __forceinline__ void init_voxels(const float* __restrict__ rhs, const float* __restrict__ mu, float mu_s,
                                            float* __restrict__ u, float* __restrict__ r, const Grid& g,
                                            const Walk& w, int m, int64_t base, float* prod) {
    for_voxels<FULL>(m, base, g, w, [&](const Vox& t, int, int, int) {
        const float a0 = -mu_at<FIELD>(mu, mu_s, t.i) * 0.0f;
        const float rn = (rhs[t.i] - a0) * (t.in ? 1.0f : 0.0f);
        if (t.live) {
            r[t.i] = rn;
            u[t.i] = 0.0f;
            prod[lds_idx(t.e)] = rn * rn;
        }
    });
}

// u = 0; r = (bz - A 0) mask, where A 0 = -mu * (+0) (solver.py:590-597); r . r.
template <bool FIELD>
__global__ void __launch_bounds__(kThreads) SW_STATIC_OCC static_init_kernel(const float* __restrict__ rhs,
                                                              const float* __restrict__ mu, float mu_s,
                                                              float* __restrict__ u, float* __restrict__ r, Grid g,
                                                              Walk w, Walk w4, int vec, int nchunks,
                                                              const SumPlan* __restrict__ plans, double* slots,
                                                              StaticState* st) {
    __shared__ float prod[kPadded];
    __shared__ float val[2 * kMaxLeaves];
    float rs;
    const bool last = chunk_sums(nchunks, g, plans, slots, st, prod, val, &rs, [&](int64_t base, int m) {
        if (m == kChunk && vec)
            init_quads<FIELD>(rhs, mu, mu_s, u, r, g, w4, base, prod);
        else if (m == kChunk)
            init_voxels<FIELD, true>(rhs, mu, mu_s, u, r, g, w, m, base, prod);
        else
            init_voxels<FIELD, false>(rhs, mu, mu_s, u, r, g, w, m, base, prod);
    });
    if (last && threadIdx.x == 0) {
        st->counter = 0u;
        st->rs_old = rs;
        st->rs0 = (kEps32 > rs) ? kEps32 : rs;  // max(rs_old, eps)
    }
}

template <bool FIRST, bool FIELD, bool FULL>
__device__ __forceinline__ void dir_voxels(const float* __restrict__ r, const float* __restrict__ p_old,
                                           float* __restrict__ p, const float* __restrict__ mu, float mu_s,
                                           const Grid& g, const Walk& w, float beta, float hx2, float hy2, float hz2,
                                           int m, int64_t base, float* prod) {
    for_voxels<FULL>(m, base, g, w, [&](const Vox& t, int x, int y, int z) {
        // p = r.copy() before the first iteration (solver.py:598), (r + beta p) mask after (solver.py:618)
        auto pn = [&](int64_t j, bool inj) {
            return FIRST ? r[j] : (r[j] + beta * p_old[j]) * (inj ? 1.0f : 0.0f);
        };
        const float pc = pn(t.i, t.in);
        // a neighbour of an interior voxel is interior unless it is on the wall of its axis; a
        // boundary voxel's taps are itself, masked as it is
        const bool nin[6] = {t.in && x + 1 < g.nx - 1, t.in && x > 1, t.in && y + 1 < g.ny - 1,
                             t.in && y > 1,            t.in && z + 1 < g.nz - 1, t.in && z > 1};
        const float lap = lap7([&](int64_t o, int q) { return pn(t.i + o, nin[q]); }, pc, t, hx2, hy2, hz2);
        const float ap = -mu_at<FIELD>(mu, mu_s, t.i) * lap;
        if (t.live) {
            prod[lds_idx(t.e)] = pc * ap;
            p[t.i] = pc;
        }
    });
}

// p (stored), p . Ap, alpha.
template <bool FIRST, bool FIELD>
__global__ void __launch_bounds__(kThreads) SW_STATIC_OCC static_dir_kernel(const float* __restrict__ r,
                                                             const float* __restrict__ p_old, float* __restrict__ p,
                                                             const float* __restrict__ mu, float mu_s, Grid g, Walk w,
                                                             Walk w4, int vec, float hx2, float hy2, float hz2,
                                                             int nchunks,
                                                             const SumPlan* __restrict__ plans, double* slots,
                                                             StaticState* st) {
    if (st->stop) return;
    __shared__ float prod[kPadded];
    __shared__ float val[2 * kMaxLeaves];
    const float beta = st->beta;
    float denom;
    const bool last = chunk_sums(nchunks, g, plans, slots, st, prod, val, &denom, [&](int64_t base, int m) {
        if (m == kChunk && vec)
            dir_quads<FIRST, FIELD>(r, p_old, p, mu, mu_s, g, w4, beta, hx2, hy2, hz2, base, prod);
        else if (m == kChunk)
            dir_voxels<FIRST, FIELD, true>(r, p_old, p, mu, mu_s, g, w, beta, hx2, hy2, hz2, m, base, prod);
        else
            dir_voxels<FIRST, FIELD, false>(r, p_old, p, mu, mu_s, g, w, beta, hx2, hy2, hz2, m, base, prod);
    });
    if (last && threadIdx.x == 0) {
        st->counter = 0u;
        if (denom <= 0.0f) {  // solver.py:601-602
            st->stop = 1;
            st->reason = SW_CG_DENOM;
        } else {
            st->alpha = st->rs_old / denom;
        }
    }
}

template <bool FIELD, bool FULL>
__device__ __forceinline__ void step_voxels(float* __restrict__ u, float* __restrict__ r, const float* __restrict__ p,
                                            const float* __restrict__ mu, float mu_s, const Grid& g, const Walk& w,
                                            float alpha, float hx2, float hy2, float hz2, int m, int64_t base,
                                            float* prod) {
    for_voxels<FULL>(m, base, g, w, [&](const Vox& t, int, int, int) {
        const float pc = p[t.i];
        const float lap = lap7([&](int64_t o, int) { return p[t.i + o]; }, pc, t, hx2, hy2, hz2);
        const float ap = -mu_at<FIELD>(mu, mu_s, t.i) * lap;
        const float mk = t.in ? 1.0f : 0.0f;
        const float un = (u[t.i] + alpha * pc) * mk;  // solver.py:604
        const float rn = (r[t.i] - alpha * ap) * mk;  // solver.py:605
        if (t.live) {
            u[t.i] = un;
            r[t.i] = rn;
            prod[lds_idx(t.e)] = rn * rn;
        }
    });
}

// u = (u + alpha p) mask, r = (r - alpha Ap) mask, r . r, stopping rules.
template <bool FIELD>
__global__ void __launch_bounds__(kThreads) SW_STATIC_OCC static_step_kernel(float* __restrict__ u, float* __restrict__ r,
                                                              const float* __restrict__ p,
                                                              const float* __restrict__ mu, float mu_s, Grid g, Walk w,
                                                              Walk w4, int vec, float hx2, float hy2, float hz2,
                                                              int nchunks,
                                                              const SumPlan* __restrict__ plans, double* slots,
                                                              StaticState* st, float* history) {
    if (st->stop) return;
    __shared__ float prod[kPadded];
    __shared__ float val[2 * kMaxLeaves];
    const float alpha = st->alpha;
    float rs_new;
    const bool last = chunk_sums(nchunks, g, plans, slots, st, prod, val, &rs_new, [&](int64_t base, int m) {
        if (m == kChunk && vec)
            step_quads<FIELD>(u, r, p, mu, mu_s, g, w4, alpha, hx2, hy2, hz2, base, prod);
        else if (m == kChunk)
            step_voxels<FIELD, true>(u, r, p, mu, mu_s, g, w, alpha, hx2, hy2, hz2, m, base, prod);
        else
            step_voxels<FIELD, false>(u, r, p, mu, mu_s, g, w, alpha, hx2, hy2, hz2, m, base, prod);
    });
    if (last && threadIdx.x == 0) {
        st->counter = 0u;
        // sqrtf and / round correctly under hipcc's defaults, as NumPy's float32
        // sqrt and divide (clang's __fsqrt_rn is the approximate native sqrt)
        const float rel = sqrtf(rs_new / st->rs0);  // solver.py:607
        const long long it = st->iters;
        st->iters = it + 1;
        st->rel = rel;
        if (history) {
            history[2 * it] = rs_new;
            history[2 * it + 1] = rel;
        }
        // rel < tol: a Python float tol compares in float32, a float64 one in float64
        const bool conv = st->tol_is_f32 ? (rel < st->tol_f) : ((double)rel < st->tol_d);
        if (conv) {
            st->stop = 1;
            st->reason = SW_CG_CONVERGED;
        } else {
            const float rs_old = st->rs_old;
            st->beta = rs_new / ((kEps32 > rs_old) ? kEps32 : rs_old);  // solver.py:617
            st->rs_old = rs_new;
            if (it + 1 >= st->max_iters) {
                st->stop = 1;
                st->reason = SW_CG_MAX_ITERS;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side

// The pairwise tree of m <= kChunk elements (NumPy's order), as a SumPlan.
struct PlanBuilder {
    struct Node {
        int left, right, height;  // children: >= 0 leaf index, < 0 internal node -(k + 1)
    };
    std::vector<std::pair<int, int>> leaves;
    std::vector<Node> nodes;

    // returns (reference, height)
    std::pair<int, int> build(int start, int m) {
        if (m <= 128) {
            leaves.push_back({start, m});
            return {(int)leaves.size() - 1, 0};
        }
        int n2 = m / 2;
        n2 -= n2 % 8;
        const auto l = build(start, n2);
        const auto r = build(start + n2, m - n2);
        const int h = 1 + (l.second > r.second ? l.second : r.second);
        nodes.push_back({l.first, r.first, h});
        return {-(int)nodes.size(), h};
    }

    bool make(int m, SumPlan* pl) {
        leaves.clear();
        nodes.clear();
        build(0, m);
        const int nl = (int)leaves.size(), ni = (int)nodes.size();
        if (nl > kMaxLeaves) return false;
        std::memset(pl, 0, sizeof(*pl));
        pl->n_leaf = nl;
        for (int j = 0; j < nl; ++j) {
            pl->leaf_start[j] = (int16_t)leaves[j].first;
            pl->leaf_len[j] = (int16_t)leaves[j].second;
        }
        // internal nodes by height (stable in creation order) get ids nl, nl + 1, ...
        std::vector<int> id(ni), order;
        int max_h = 0;
        for (const Node& nd : nodes) max_h = nd.height > max_h ? nd.height : max_h;
        if (max_h > kMaxLevels) return false;
        for (int h = 1; h <= max_h; ++h) {
            for (int k = 0; k < ni; ++k)
                if (nodes[k].height == h) {
                    id[k] = nl + (int)order.size();
                    order.push_back(k);
                }
            pl->level_end[h - 1] = (int16_t)order.size();
        }
        pl->n_level = max_h;
        auto ref = [&](int v) { return v >= 0 ? v : id[-v - 1]; };
        for (int q = 0; q < ni; ++q) {
            pl->node_l[q] = (int16_t)ref(nodes[order[q]].left);
            pl->node_r[q] = (int16_t)ref(nodes[order[q]].right);
        }
        return true;
    }
};

int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// Persistent blocks: a few per CU, each walking several chunks.
#ifndef SW_STATIC_BLOCKS_PER_CU
#define SW_STATIC_BLOCKS_PER_CU 2
#endif

int static_grid(int nchunks) {
    const int cap = device_cu_count() * SW_STATIC_BLOCKS_PER_CU;
    return nchunks < cap ? nchunks : cap;
}

int64_t n_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

int64_t static_workspace(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t n = nx * ny * nz;
    return 3 * align256(n * 4) + align256(n_chunks(n) * 8) + align256(2 * (int64_t)sizeof(SumPlan)) +
           align256(sizeof(StaticState));
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int64_t sw_static_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    return static_workspace(nx, ny, nz);
}

int sw_elastic_solve_static(const float* bz, const float* mu, double mu_scalar, int64_t nx, int64_t ny, int64_t nz,
                            double dx, double dy, double dz, double tol, int tol_is_f32, int64_t max_iters, float* u,
                            void* workspace, int64_t workspace_bytes_given, float* history, sw_static_info* info,
                            void* stream) {
    int rc = check_grid(nx, ny, nz);
    if (rc) return rc;
    SW_REQUIRE(bz && u, SW_ERR_ARG, "null bz or u");
    SW_REQUIRE(max_iters >= 0, SW_ERR_ARG, "max_iters must be >= 0");
    SW_REQUIRE(dx > 0 && dy > 0 && dz > 0, SW_ERR_ARG, "grid spacing must be positive");
    const int64_t need = static_workspace(nx, ny, nz);
    SW_REQUIRE(workspace && workspace_bytes_given >= need, SW_ERR_WORKSPACE,
               "workspace of %lld bytes is smaller than the %lld needed", (long long)workspace_bytes_given,
               (long long)need);
    SW_REQUIRE(((uintptr_t)workspace & 255) == 0, SW_ERR_ARG, "workspace must be 256-byte aligned");

    const int64_t n = nx * ny * nz;
    const int64_t nc64 = n_chunks(n);
    SW_REQUIRE(nc64 < (int64_t(1) << 30), SW_ERR_ARG, "grid too large");
    const int nc = (int)nc64;
    char* wp = (char*)workspace;
    float* r = (float*)wp;
    wp += align256(n * 4);
    float* pa = (float*)wp;
    wp += align256(n * 4);
    float* pb = (float*)wp;
    wp += align256(n * 4);
    double* slots = (double*)wp;
    wp += align256(nc64 * 8);
    SumPlan* plans = (SumPlan*)wp;
    wp += align256(2 * (int64_t)sizeof(SumPlan));
    StaticState* st = (StaticState*)wp;

    // plans[0]: a full chunk; plans[1]: the last, partial one (if any)
    SumPlan hp[2];
    PlanBuilder pb_host;
    SW_REQUIRE(pb_host.make(kChunk, &hp[0]), SW_ERR_ARG, "pairwise plan");
    const int tail = (int)(n % kChunk);
    SW_REQUIRE(pb_host.make(tail ? tail : kChunk, &hp[1]), SW_ERR_ARG, "pairwise plan");

    hipStream_t s = (hipStream_t)stream;
    SW_TRY_HIP(hipMemcpyAsync(plans, hp, sizeof(hp), hipMemcpyHostToDevice, s));
    Grid g{(int)nx, (int)ny, (int)nz};
    const Walk w{(int)(kThreads / (ny * nz)), (int)((kThreads / nz) % ny), (int)(kThreads % nz)};
    const Walk w4{(int)(kQuad / (ny * nz)), (int)((kQuad / nz) % ny), (int)(kQuad % nz)};
    auto a16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
    const int vec = nz % 4 == 0 && a16(bz) && a16(u) && (!mu || a16(mu));  // workspace arrays are 256-aligned
    const float hx2 = (float)(dx * dx), hy2 = (float)(dy * dy), hz2 = (float)(dz * dz);  // dX**2 in float32
    const float mu_s = (float)mu_scalar;

    hipLaunchKernelGGL(static_reset_kernel, dim3(1), dim3(1), 0, s, st, tol, tol_is_f32, (long long)max_iters);
    SW_LAUNCH_CHECK();
    const bool field = mu != nullptr;
    const int nb = static_grid(nc);
    hipLaunchKernelGGL(field ? static_init_kernel<true> : static_init_kernel<false>, dim3(nb), dim3(kThreads), 0, s,
                       bz, mu, mu_s, u, r, g, w, w4, vec, nc, plans, slots, st);
    SW_LAUNCH_CHECK();

    StaticState hs;
    SW_TRY_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
    SW_TRY_HIP(hipStreamSynchronize(s));
    int64_t launched = 0, chunk = 8;
    float* p_old = pa;
    float* p_new = pb;
    while (!hs.stop && launched < max_iters) {
        const int64_t k = (max_iters - launched < chunk) ? (max_iters - launched) : chunk;
        for (int64_t j = 0; j < k; ++j) {
            // iteration 0 takes p = r
            const bool first = launched + j == 0;
            auto dir = first ? (field ? static_dir_kernel<true, true> : static_dir_kernel<true, false>)
                             : (field ? static_dir_kernel<false, true> : static_dir_kernel<false, false>);
            hipLaunchKernelGGL(dir, dim3(nb), dim3(kThreads), 0, s, r, p_old, p_new, mu, mu_s, g, w, w4, vec, hx2,
                               hy2, hz2, nc, plans, slots, st);
            hipLaunchKernelGGL(field ? static_step_kernel<true> : static_step_kernel<false>, dim3(nb),
                               dim3(kThreads), 0, s, u, r, p_new, mu, mu_s, g, w, w4, vec, hx2, hy2, hz2, nc, plans,
                               slots, st, history);
            float* t = p_old;
            p_old = p_new;
            p_new = t;
        }
        SW_LAUNCH_CHECK();
        launched += k;
        SW_TRY_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
        SW_TRY_HIP(hipStreamSynchronize(s));
        if (chunk < 64) chunk *= 2;
    }
    if (info) {
        info->iters = hs.iters;
        info->relres = hs.rel;
        info->reason = hs.reason;
        info->pad = 0;
    }
    return SW_OK;
}

}  // extern "C"
