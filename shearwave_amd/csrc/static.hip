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
    int stop, reason, first, tol_is_f32;
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

// Chunks of an XCD (blocks b with b % 8 == xcd) are contiguous.
__device__ __forceinline__ int chunk_of_block(int nb) {
    const int b = blockIdx.x, q = nb >> 3, rem = nb & 7, xcd = b & 7;
    return xcd * q + min(xcd, rem) + (b >> 3);
}

// ((c+ - 2c) + c-) / h^2 per axis, summed x, y, z (laplacian_center, solver.py:686-695).
template <typename V>
__device__ __forceinline__ float lap7(V v, float c, int64_t P, int64_t R, float hx2, float hy2, float hz2) {
    const float c2 = 2.0f * c;
    const float tx = ((v(P, 0) - c2) + v(-P, 1)) / hx2;
    const float ty = ((v(R, 2) - c2) + v(-R, 3)) / hy2;
    const float tz = ((v((int64_t)1, 4) - c2) + v((int64_t)-1, 5)) / hz2;
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
#pragma unroll 8
            for (int q = 0; q < cnt; ++q) tot = tot + buf[q];
        }
        __syncthreads();
    }
    return tot;
}

// Publish this chunk's sum; in the block that publishes last, return true and
// the total in thread 0.
__device__ __forceinline__ bool finish_sum(float s, int chunk, int nchunks, double* slots, StaticState* st,
                                           float* buf, float* total) {
    if (!publish_slot(slots, nullptr, chunk, (double)s, 0.0, &st->counter, (unsigned)nchunks)) return false;
    *total = ordered_total(slots, nchunks, buf);
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
    st->first = 1;
    st->counter = 0u;
}

// u = 0; r = (bz - A 0) mask, where A 0 = -mu * (+0) (solver.py:590-597); r . r.
__global__ void __launch_bounds__(kThreads) static_init_kernel(const float* __restrict__ rhs,
                                                              const float* __restrict__ mu, float mu_s,
                                                              float* __restrict__ u, float* __restrict__ r, Grid g,
                                                              Walk w, int nchunks, const SumPlan* __restrict__ plans,
                                                              double* slots, StaticState* st) {
    __shared__ float prod[kPadded];
    __shared__ float val[2 * kMaxLeaves];
    const int c = chunk_of_block(nchunks);
    const int64_t base = (int64_t)c * kChunk;
    const int m = (int)min<int64_t>(kChunk, g.size() - base);
    int x, y, z;
    coords(base + threadIdx.x, g, x, y, z);
    for (int k = 0; k < kPerThread; ++k) {
        const int e = threadIdx.x + k * kThreads;
        if (e >= m) break;
        const int64_t i = base + e;
        const float muv = mu ? mu[i] : mu_s;
        const float a0 = -muv * 0.0f;
        const float rn = (rhs[i] - a0) * (g.interior(x, y, z) ? 1.0f : 0.0f);
        r[i] = rn;
        u[i] = 0.0f;
        prod[lds_idx(e)] = rn * rn;
        advance(x, y, z, g, w);
    }
    __syncthreads();
    const float s = plan_sum(prod, val, plans[m == kChunk ? 0 : 1]);
    float rs;
    if (finish_sum(s, c, nchunks, slots, st, prod, &rs) && threadIdx.x == 0) {
        st->counter = 0u;
        st->rs_old = rs;
        st->rs0 = (kEps32 > rs) ? kEps32 : rs;  // max(rs_old, eps)
    }
}

// p = r (first iteration) or (r + beta p_old) mask; stores p; p . Ap; alpha.
__global__ void __launch_bounds__(kThreads) static_dir_kernel(const float* __restrict__ r,
                                                             const float* __restrict__ p_old, float* __restrict__ p,
                                                             const float* __restrict__ mu, float mu_s, Grid g, Walk w,
                                                             float hx2, float hy2, float hz2, int nchunks,
                                                             const SumPlan* __restrict__ plans, double* slots,
                                                             StaticState* st) {
    if (st->stop) return;
    __shared__ float prod[kPadded];
    __shared__ float val[2 * kMaxLeaves];
    const int c = chunk_of_block(nchunks);
    const int64_t base = (int64_t)c * kChunk;
    const int m = (int)min<int64_t>(kChunk, g.size() - base);
    const bool first = st->first != 0;
    const float beta = st->beta;
    const int64_t P = g.plane(), R = g.nz;
    int x, y, z;
    coords(base + threadIdx.x, g, x, y, z);
    for (int k = 0; k < kPerThread; ++k) {
        const int e = threadIdx.x + k * kThreads;
        if (e >= m) break;
        const int64_t i = base + e;
        const bool in = g.interior(x, y, z);
        auto pn = [&](int64_t j, bool inj) {
            return first ? r[j] : (r[j] + beta * p_old[j]) * (inj ? 1.0f : 0.0f);
        };
        const float pc = pn(i, in);
        float lap = 0.0f;
        if (in) {
            // a neighbour of an interior voxel is interior unless it is on the wall of its axis
            const bool nin[6] = {x + 1 < g.nx - 1, x > 1, y + 1 < g.ny - 1, y > 1, z + 1 < g.nz - 1, z > 1};
            lap = lap7([&](int64_t o, int q) { return pn(i + o, nin[q]); }, pc, P, R, hx2, hy2, hz2);
        }
        const float muv = mu ? mu[i] : mu_s;
        const float ap = -muv * lap;
        prod[lds_idx(e)] = pc * ap;
        p[i] = pc;
        advance(x, y, z, g, w);
    }
    __syncthreads();
    const float s = plan_sum(prod, val, plans[m == kChunk ? 0 : 1]);
    float denom;
    if (finish_sum(s, c, nchunks, slots, st, prod, &denom) && threadIdx.x == 0) {
        st->counter = 0u;
        if (denom <= 0.0f) {  // solver.py:601-602
            st->stop = 1;
            st->reason = SW_CG_DENOM;
        } else {
            st->alpha = st->rs_old / denom;
        }
    }
}

// u = (u + alpha p) mask, r = (r - alpha Ap) mask, r . r, stopping rules.
__global__ void __launch_bounds__(kThreads) static_step_kernel(float* __restrict__ u, float* __restrict__ r,
                                                              const float* __restrict__ p,
                                                              const float* __restrict__ mu, float mu_s, Grid g, Walk w,
                                                              float hx2, float hy2, float hz2, int nchunks,
                                                              const SumPlan* __restrict__ plans, double* slots,
                                                              StaticState* st, float* history) {
    if (st->stop) return;
    __shared__ float prod[kPadded];
    __shared__ float val[2 * kMaxLeaves];
    const int c = chunk_of_block(nchunks);
    const int64_t base = (int64_t)c * kChunk;
    const int m = (int)min<int64_t>(kChunk, g.size() - base);
    const float alpha = st->alpha;
    const int64_t P = g.plane(), R = g.nz;
    int x, y, z;
    coords(base + threadIdx.x, g, x, y, z);
    for (int k = 0; k < kPerThread; ++k) {
        const int e = threadIdx.x + k * kThreads;
        if (e >= m) break;
        const int64_t i = base + e;
        const bool in = g.interior(x, y, z);
        const float pc = p[i];
        float lap = 0.0f;
        if (in) lap = lap7([&](int64_t o, int) { return p[i + o]; }, pc, P, R, hx2, hy2, hz2);
        const float muv = mu ? mu[i] : mu_s;
        const float ap = -muv * lap;
        const float mk = in ? 1.0f : 0.0f;
        const float un = (u[i] + alpha * pc) * mk;
        const float rn = (r[i] - alpha * ap) * mk;
        u[i] = un;
        r[i] = rn;
        prod[lds_idx(e)] = rn * rn;
        advance(x, y, z, g, w);
    }
    __syncthreads();
    const float s = plan_sum(prod, val, plans[m == kChunk ? 0 : 1]);
    float rs_new;
    if (finish_sum(s, c, nchunks, slots, st, prod, &rs_new) && threadIdx.x == 0) {
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
            st->first = 0;
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
    const float hx2 = (float)(dx * dx), hy2 = (float)(dy * dy), hz2 = (float)(dz * dz);  // dX**2 in float32
    const float mu_s = (float)mu_scalar;

    hipLaunchKernelGGL(static_reset_kernel, dim3(1), dim3(1), 0, s, st, tol, tol_is_f32, (long long)max_iters);
    SW_LAUNCH_CHECK();
    hipLaunchKernelGGL(static_init_kernel, dim3(nc), dim3(kThreads), 0, s, bz, mu, mu_s, u, r, g, w, nc, plans,
                       slots, st);
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
            hipLaunchKernelGGL(static_dir_kernel, dim3(nc), dim3(kThreads), 0, s, r, p_old, p_new, mu, mu_s, g, w,
                               hx2, hy2, hz2, nc, plans, slots, st);
            hipLaunchKernelGGL(static_step_kernel, dim3(nc), dim3(kThreads), 0, s, u, r, p_new, mu, mu_s, g, w, hx2,
                               hy2, hz2, nc, plans, slots, st, history);
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
