// Batched solves on gfx950: many shear_fdtd solves of uneven grid sizes and
// material maps advance in ONE persistent launch per pass (two steps, or one,
// per solve as its pass plan says), and their projections run as one batched
// CG.
//
// Work items are (solve, z tile, y tile, x chunk) quadruples. A persistent
// grid (one wave of resident blocks) walks the item list; each XCD takes a
// contiguous range of items, so tiles that share halos share an L2. Each item
// runs exactly the tile code of the single-solve kernels (march_step.h,
// march_cg.h), and every reduction of a solve goes to that solve's own slots
// in a fixed order, so a solve's result does not depend on which other solves
// share the batch.
#include <algorithm>
#include <vector>

#include "march2_step.h"
#include "march_cg.h"
#include "march_step.h"
#include "sw_common.h"

namespace sw {
namespace {

// Contiguous item range of this block's XCD; k runs lo + j, lo + j + stride, ...
struct ItemWalk {
    int64_t lo, hi, first, stride;
};

__device__ inline ItemWalk item_walk(int64_t n_items) {
    ItemWalk w;
    const unsigned nb = gridDim.x;
    if ((nb & 7u) == 0u) {
        const int64_t xcd = blockIdx.x & 7u, per = nb >> 3, j = blockIdx.x >> 3;
        w.lo = xcd * n_items / 8;
        w.hi = (xcd + 1) * n_items / 8;
        w.first = w.lo + j;
        w.stride = per;
    } else {
        w.lo = 0;
        w.hi = n_items;
        w.first = blockIdx.x;
        w.stride = nb;
    }
    return w;
}

// Dynamic item claims: each block takes the next item with one atomic on
// claim[0] and runs body(k) until the list is exhausted, so blocks that drew
// short items take more of them and a launch ends when its work does (the host
// lists the items longest first). The last block to finish resets claim[0]
// and claim[1] to zero for the next launch, so no memset is needed (and none is
// recorded into a graph). body must be uniform over the block.
template <typename F>
__device__ inline void claim_items(unsigned long long* claim, int64_t n_items, F&& body) {
    __shared__ long long claimed;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    for (;;) {
        if (tid == 0) claimed = (long long)atomicAdd(claim, 1ull);
        __syncthreads();  // the claim is visible; the previous item's LDS use is done
        const long long k = claimed;
        __syncthreads();  // every thread has read the claim before thread 0 makes the next one
        if (k >= n_items) break;
        body(k);
    }
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(claim + 1, 1ull) == (unsigned long long)gridDim.x - 1ull) {
            claim[0] = 0ull;  // every block has made its last claim
            claim[1] = 0ull;
            __threadfence();
        }
    }
}

// Claims per XCD: the item list is cut into 8 equal ranges, one per XCD (a
// block's XCD is blockIdx % 8: the hardware deals blocks round-robin), and a
// block claims items of its own XCD's range first, so tiles listed next to
// each other (neighbours, when the list is in spatial order) run on one XCD
// at about the same time and share the lines at their edges in its L2. A
// block whose range is exhausted claims from the others'. ctr[0..7] are the
// ranges' counters and ctr[8] counts finished blocks; the last block resets
// them all.
template <typename F>
__device__ inline void claim_items_xcd(unsigned long long* ctr, int64_t n_items, F&& body) {
    __shared__ long long claimed;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nr = (gridDim.x & 7u) == 0u ? 8 : 1;
    const int own = nr == 8 ? (int)(blockIdx.x & 7u) : 0;
    for (int t = 0; t < nr; ++t) {
        const int j = (own + t) % nr;
        const long long lo = (long long)j * n_items / nr, hi = (long long)(j + 1) * n_items / nr;
        for (;;) {
            if (tid == 0) claimed = lo + (long long)atomicAdd(ctr + j, 1ull);
            __syncthreads();  // the claim is visible; the previous item's LDS use is done
            const long long k = claimed;
            __syncthreads();  // every thread has read the claim before thread 0 makes the next one
            if (k >= hi) break;
            body(k);
        }
    }
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(ctr + 8, 1ull) == (unsigned long long)gridDim.x - 1ull) {
            for (int j = 0; j < 9; ++j) ctr[j] = 0ull;  // every block has made its last claim
            __threadfence();
        }
    }
}

// Tile rows of the batched step: BY output rows (4, 8 or 12, one launch of
// one shape; the autotuner picks it with the x-chunk cap); the block has
// BY + 3 rows of threads so that a two-step pass (march2_step.h) fits, and
// single steps leave the three extra rows to the barriers. The 4-row tile
// prefetches one plane ahead so that two 7-wave blocks share a CU; the 8-row
// tile two planes (one 11-wave block per CU), the 12-row tile one.
constexpr int kBatchBYDefault = 8;

// Waves per SIMD the kernel must fit (launch bound): two 7-wave blocks per CU
// for the 4-row tile, else whatever one block needs.
template <int BY>
constexpr int batch_min_waves_per_simd() {
    return BY == 4 ? 4 : (BY + 3 + 3) / 4;
}

template <int BY, bool FORCE>
constexpr int batch_step2_dist() {
    return BY == 4 ? 1 : 0;  // 0: march2_tile's default
}

// LDS of one block: the two-step tile's planes, or the single-step tile's.
template <int BY>
union BatchLds {
    Step2Lds<float, true, BY> two_visc;
    Step2Lds<float, false, BY> two_inviscid;
    float one[2][6][BY + 2][66];
};

// pass_state[0] is the global pass; items are claimed on pass_state[1..9].
template <int BY>
__global__ void __launch_bounds__(64 * (BY + 3), batch_min_waves_per_simd<BY>()) batch_step_kernel(const sw_batch_solve* __restrict__ solves,
                                                                   const int4* __restrict__ items, int64_t n_items,
                                                                   long long* __restrict__ pass_state) {
    constexpr int R = BY + 3;
    __shared__ BatchLds<BY> lds;
    const long long gp = pass_state[0];
    claim_items_xcd((unsigned long long*)(pass_state + 1), n_items, [&](long long k) {
        const int4 it = items[k];
        const sw_batch_solve& S = solves[it.x];
        const long long pass = gp - S.start;
        if (pass < 0 || pass >= S.n_passes) return;  // uniform in the block
        long long local = pass;
        bool two = false;
        if (S.plan) {
            const int e = S.plan[pass];
            local = e >> 1;
            two = e & 1;
        }
        const int par = (int)(pass & 1);
        StepArgs<float> a;
        for (int c = 0; c < 3; ++c) {
            a.A[c] = S.u[par][c];
            a.B[c] = S.d[par][c];
            a.Ao[c] = S.u[par ^ 1][c];
            a.Bo[c] = S.d[par ^ 1][c];
            a.f[c] = S.force[c];
        }
        a.m = S.mu_over_rho;
        a.e = S.eta_over_rho;
        a.g = Grid{S.nx, S.ny, S.nz};
        a.cx = S.cx, a.cy = S.cy, a.cz = S.cz;
        a.vx = S.vx, a.vy = S.vy, a.vz = S.vz;
        a.damp = S.damp;
        const bool forcing = S.env && S.force[0];
        a.fs = forcing ? (float)S.env[local] : 0.f;
        a.fs2 = (forcing && two) ? (float)S.env[local + 1] : 0.f;
        const int x0 = it.w * S.xchunk;
        const int x1 = min(x0 + S.xchunk, S.nx);
        const int y0 = it.z * BY, z0 = it.y * 64;
        const bool visc = a.e != nullptr, force = a.fs != 0.f || a.fs2 != 0.f;
        if (two) {
            if (visc) {
                if (force)
                    march2_tile<float, true, true, BY, false, batch_step2_dist<BY, true>()>(a, x0, x1, y0, z0,
                                                                                   lds.two_visc);
                else
                    march2_tile<float, true, false, BY, false, batch_step2_dist<BY, false>()>(a, x0, x1, y0, z0,
                                                                                   lds.two_visc);
            } else {
                if (force)
                    march2_tile<float, false, true, BY, false, batch_step2_dist<BY, true>()>(a, x0, x1, y0, z0,
                                                                                   lds.two_inviscid);
                else
                    march2_tile<float, false, false, BY, false, batch_step2_dist<BY, false>()>(a, x0, x1, y0, z0,
                                                                                   lds.two_inviscid);
            }
        } else if (visc) {
            if (force)
                march_tile<float, true, true, true, BY, 64, 6, R>(a, x0, x1, y0, z0, lds.one);
            else
                march_tile<float, true, true, false, BY, 64, 6, R>(a, x0, x1, y0, z0, lds.one);
        } else {
            if (force)
                march_tile<float, true, false, true, BY, 64, 6, R>(a, x0, x1, y0, z0, lds.one);
            else
                march_tile<float, true, false, false, BY, 64, 6, R>(a, x0, x1, y0, z0, lds.one);
        }
    });
}

__global__ void advance_kernel(long long* counter) { *counter += 1; }

// ------------------------------------------------------------ batched CG

// states[n] holds the dir and step kernels' work-claim counters.
__global__ void batch_cg_reset_kernel(const sw_batch_cg* solves, int n, CgState* states) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) cg_reset(states + s, solves[s].tol, (long long)solves[s].max_iters, solves[s].tol);
    if (s == n) {
        unsigned long long* claim = (unsigned long long*)(states + n);
        claim[0] = claim[1] = 0ull;
    }
}

__device__ inline unsigned long long* cg_claims(CgState* states, int n_solves) {
    return (unsigned long long*)(states + n_solves);
}

__device__ inline CgTile batch_cg_tile(const sw_batch_cg& S, int4 it) {
    const Grid g{S.nx, S.ny, S.nz};
    const int x0 = it.w * S.xchunk;
    return cg_tile(g, x0, min(x0 + S.xchunk, S.nx), it.z * kCgBY, it.y * 64);
}

// A solve's partial slots: one per (x sub-chunk of kCgSub planes, y tile,
// z tile), in that order. An item of xchunk planes (a multiple of kCgSub, or
// all of nx) publishes one slot per sub-chunk it covers, each summed over the
// same planes in the same order as an item of kCgSub planes would, so a
// solve's dot products do not depend on the item length.
__device__ inline int cg_n_slots(const sw_batch_cg& S) {
    return ((S.nx + kCgSub - 1) / kCgSub) * ((S.ny + kCgBY - 1) / kCgBY) * ((S.nz + 63) / 64);
}

__device__ inline int cg_slot(const sw_batch_cg& S, int4 it, int j) {
    const int nty = (S.ny + kCgBY - 1) / kCgBY, ntz = (S.nz + 63) / 64;
    return ((it.w * S.xchunk / kCgSub + j) * nty + it.z) * ntz + it.y;
}

__global__ void __launch_bounds__(64 * kCgBY) batch_cg_init_kernel(const sw_batch_cg* __restrict__ solves, CgState* states,
                                                                   const int4* __restrict__ items, int64_t n_items) {
    __shared__ double red[16];
    const ItemWalk w = item_walk(n_items);
    for (int64_t k = w.first; k < w.hi; k += w.stride) {
        const int4 it = items[k];
        const sw_batch_cg& S = solves[it.x];
        CgState* st = states + it.x;
        const Grid g{S.nx, S.ny, S.nz};
        const CgTile t = batch_cg_tile(S, it);
        const int n_slots = cg_n_slots(S);
        for (int j = 0; t.x0 + j * kCgSub < t.x1; ++j) {
            CgTile ts = t;
            ts.x0 = t.x0 + j * kCgSub;
            ts.x1 = min(ts.x0 + kCgSub, t.x1);
            double s = 0.0, mx = 0.0;
            cg_init_tile<float, float>(S.f[0], S.f[1], S.f[2], S.phi, S.r, S.p0, g, 0.5 / S.dx, 0.5 / S.dy,
                                       0.5 / S.dz, ts, s, mx);
            s = block_sum(s, red);
            mx = block_max(mx, red);
            if (publish_slot(S.part, S.part + n_slots, cg_slot(S, it, j), s, mx, &st->counter, n_slots)) {
                double tot, m;
                reduce_slots(S.part, S.part + n_slots, n_slots, red, &tot, &m);
                if (threadIdx.x == 0 && threadIdx.y == 0) {
                    const int stop = st->stop, reason = st->reason;  // max_iters == 0 keeps its stop
                    cg_after_init(st, tot, m);
                    if (stop && !st->stop) {
                        st->stop = stop;
                        st->reason = reason;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// p ping-pongs between p0 and p1 by the solve's iteration parity. The batch
// keeps r and p in float32 (phi, dot products and residual norm in float64).
__global__ void __launch_bounds__(64 * kCgBY) batch_cg_dir_kernel(const sw_batch_cg* __restrict__ solves, CgState* states,
                                                                  int n_solves, const int4* __restrict__ items,
                                                                  int64_t n_items) {
    __shared__ CgLdsT<float> lds;
    __shared__ double red[16];
    claim_items(cg_claims(states, n_solves), n_items, [&](long long k) {
        const int4 it = items[k];
        const sw_batch_cg& S = solves[it.x];
        CgState* st = states + it.x;
        if (st->stop) return;
        const Grid g{S.nx, S.ny, S.nz};
        const CgTile t = batch_cg_tile(S, it);
        const bool odd = st->iters & 1;
        const float* p_old = odd ? S.p1 : S.p0;
        float* p_new = odd ? S.p0 : S.p1;
        const int n_slots = cg_n_slots(S);
        // one slot per sub-chunk, stored as it ends and counted once per item:
        // the block that completes a solve's slots has passed every other item
        // of the solve, which read beta before marching
        auto sink = [&](double& share, int j) {
            const double b = block_sum(share, red);
            share = 0.0;
            if (threadIdx.x == 0 && threadIdx.y == 0) S.part[cg_slot(S, it, j)] = b;
        };
        cg_dir_tile<float>(S.r, p_old, p_new, g, st->beta, 0.25 / (S.dx * S.dx), 0.25 / (S.dy * S.dy),
                           0.25 / (S.dz * S.dz), t, lds, sink);
        if (publish_count(&st->counter, (t.x1 - t.x0 + kCgSub - 1) / kCgSub, n_slots)) {
            double denom, unused;
            reduce_slots(S.part, nullptr, n_slots, red, &denom, &unused);
            if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_dir(st, denom);
        }
    });
}

__global__ void __launch_bounds__(64 * kCgBY) batch_cg_step_kernel(const sw_batch_cg* __restrict__ solves, CgState* states,
                                                                   int n_solves, const int4* __restrict__ items,
                                                                   int64_t n_items) {
    __shared__ CgLdsT<float> lds;
    __shared__ double red[16];
    claim_items(cg_claims(states, n_solves), n_items, [&](long long k) {
        const int4 it = items[k];
        const sw_batch_cg& S = solves[it.x];
        CgState* st = states + it.x;
        if (st->stop) return;
        const Grid g{S.nx, S.ny, S.nz};
        const CgTile t = batch_cg_tile(S, it);
        // iteration k = iters + 1 wrote p into p1 (k odd) or p0 (k even); phi is
        // updated on even iterations for both (march_cg.h), so p_prev is the other buffer
        const bool odd = st->iters & 1;
        const float* p = odd ? S.p0 : S.p1;
        const float* p_prev = odd ? S.p1 : S.p0;
        const double qx = 0.25 / (S.dx * S.dx), qy = 0.25 / (S.dy * S.dy), qz = 0.25 / (S.dz * S.dz);
        const int n_slots = cg_n_slots(S);
        auto sink = [&](double& share, int j) {
            const double b = block_sum(share, red);
            share = 0.0;
            if (threadIdx.x == 0 && threadIdx.y == 0) S.part[cg_slot(S, it, j)] = b;
        };
        const double alpha = st->alpha, alpha_prev = st->alpha_prev;
        if (odd)
            cg_step_tile<float, kPhiPair>(S.phi, S.r, p, p_prev, g, alpha, alpha_prev, qx, qy, qz, t, lds, sink);
        else
            cg_step_tile<float, kPhiSkip>(S.phi, S.r, p, p_prev, g, alpha, alpha_prev, qx, qy, qz, t, lds, sink);
        if (publish_count(&st->counter, (t.x1 - t.x0 + kCgSub - 1) / kCgSub, n_slots)) {
            double rs_new, unused;
            reduce_slots(S.part, nullptr, n_slots, red, &rs_new, &unused);
            if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_step(st, rs_new);
        }
    });
}

// The deferred phi term of each solve that stopped after an odd iteration.
__global__ void __launch_bounds__(64 * kCgBY) batch_cg_flush_kernel(const sw_batch_cg* __restrict__ solves,
                                                                    const CgState* states,
                                                                    const int4* __restrict__ items, int64_t n_items) {
    const ItemWalk w = item_walk(n_items);
    for (int64_t k = w.first; k < w.hi; k += w.stride) {
        const int4 it = items[k];
        const sw_batch_cg& S = solves[it.x];
        const CgState* st = states + it.x;
        if ((st->iters & 1) == 0) continue;
        const Grid g{S.nx, S.ny, S.nz};
        cg_flush_phi_tile<float>(S.phi, S.p1, g, st->alpha, batch_cg_tile(S, it));
    }
}

__global__ void __launch_bounds__(64 * kCgBY) batch_cg_grad_sub_kernel(const sw_batch_cg* __restrict__ solves,
                                                                       const int4* __restrict__ items,
                                                                       int64_t n_items) {
    const ItemWalk w = item_walk(n_items);
    for (int64_t k = w.first; k < w.hi; k += w.stride) {
        const int4 it = items[k];
        const sw_batch_cg& S = solves[it.x];
        const Grid g{S.nx, S.ny, S.nz};
        const CgTile t = batch_cg_tile(S, it);
        cg_grad_sub_tile<float>(S.f[0], S.f[1], S.f[2], S.phi, g, 0.5 / S.dx, 0.5 / S.dy, 0.5 / S.dz, S.out[0],
                                S.out[1], S.out[2], S.inv_rho, S.scale, t);
    }
}

template <int BY>
unsigned persistent_step_blocks(int requested) {
    if (requested > 0) return (unsigned)requested;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_step_kernel<BY>, 64 * (BY + 3), 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return (unsigned)(per_cu * device_cu_count());  // a multiple of 8 on an 8-XCD device
}

template <int BY>
void launch_batch_passes_rows(const sw_batch_solve* solves, const int32_t* items, int64_t n_items, int64_t* gpass,
                              int blocks, int64_t n_passes, hipStream_t s) {
    const unsigned nb = persistent_step_blocks<BY>(blocks);
    for (int64_t k = 0; k < n_passes; ++k) {
        hipLaunchKernelGGL(batch_step_kernel<BY>, dim3(nb), dim3(64, BY + 3), 0, s, solves, (const int4*)items,
                           n_items, (long long*)gpass);
        hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, s, (long long*)gpass);
    }
}

bool batch_rows_supported(int tile_y) { return tile_y == 4 || tile_y == 8 || tile_y == 12; }

int launch_batch_passes(const sw_batch_solve* solves, const int32_t* items, int64_t n_items, int64_t* gpass,
                        int tile_y, int blocks, int64_t n_passes, hipStream_t s) {
    switch (tile_y) {
        case 4:
            launch_batch_passes_rows<4>(solves, items, n_items, gpass, blocks, n_passes, s);
            break;
        case 12:
            launch_batch_passes_rows<12>(solves, items, n_items, gpass, blocks, n_passes, s);
            break;
        default:
            launch_batch_passes_rows<kBatchBYDefault>(solves, items, n_items, gpass, blocks, n_passes, s);
            break;
    }
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // namespace
}  // namespace sw

struct sw_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

using namespace sw;

extern "C" {

int sw_batch_step(const sw_batch_solve* solves, const int32_t* items, int64_t n_items, int64_t* gpass, int32_t tile_y,
                  int32_t blocks, int64_t n_passes, void* stream) {
    SW_REQUIRE(solves && gpass && (items || n_items == 0), SW_ERR_ARG, "null pointer");
    SW_REQUIRE(batch_rows_supported(tile_y), SW_ERR_ARG, "tile_y %d must be 4, 8 or 12", tile_y);
    SW_REQUIRE(n_items >= 0 && n_passes >= 0 && blocks >= 0, SW_ERR_ARG, "negative size");
    return launch_batch_passes(solves, items, n_items, gpass, tile_y, blocks, n_passes, (hipStream_t)stream);
}

int sw_batch_graph_create(const sw_batch_solve* solves, const int32_t* items, int64_t n_items, int64_t* gpass,
                          int32_t tile_y, int32_t blocks, int64_t n_passes, sw_graph** graph) {
    SW_REQUIRE(graph, SW_ERR_ARG, "null graph handle");
    SW_REQUIRE(solves && gpass && (items || n_items == 0), SW_ERR_ARG, "null pointer");
    SW_REQUIRE(batch_rows_supported(tile_y), SW_ERR_ARG, "tile_y %d must be 4, 8 or 12", tile_y);
    SW_REQUIRE(n_items >= 0 && n_passes >= 1 && blocks >= 0, SW_ERR_ARG, "bad size");
    *graph = nullptr;
    hipStream_t cap;
    SW_TRY_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cap);
        SW_TRY_HIP(e);
    }
    const int rc = launch_batch_passes(solves, items, n_items, gpass, tile_y, blocks, n_passes, cap);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (rc != SW_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SW_TRY_HIP(e);
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        SW_TRY_HIP(e);
    }
    *graph = new sw_graph{g, ex};
    return SW_OK;
}

int sw_graph_launch(sw_graph* graph, void* stream) {
    SW_REQUIRE(graph, SW_ERR_ARG, "null graph");
    SW_TRY_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return SW_OK;
}

int sw_graph_destroy(sw_graph* graph) {
    if (!graph) return SW_OK;
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    delete graph;
    return SW_OK;
}

int64_t sw_cg_state_bytes(void) { return (int64_t)sizeof(CgState); }

namespace {
// The descriptors and items (device memory), checked on the host before any
// launch: a solve whose slots were not each covered once would never complete
// its reductions, and the host loop would wait on it.
int check_batch_cg(const sw_batch_cg* solves, int32_t n_solves, const int32_t* items, int64_t n_items,
                   hipStream_t s) {
    std::vector<sw_batch_cg> d(n_solves);
    std::vector<int4> it(n_items);
    SW_TRY_HIP(hipMemcpyAsync(d.data(), solves, sizeof(sw_batch_cg) * n_solves, hipMemcpyDeviceToHost, s));
    SW_TRY_HIP(hipMemcpyAsync(it.data(), items, sizeof(int4) * n_items, hipMemcpyDeviceToHost, s));
    SW_TRY_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> base(n_solves + 1, 0);  // each solve's slots in one covered-once table
    for (int k = 0; k < n_solves; ++k) {
        const sw_batch_cg& S = d[k];
        SW_REQUIRE(S.nx >= 1 && S.ny >= 1 && S.nz >= 1 && (int64_t)S.ny * S.nz < kMaxMarchPlane, SW_ERR_ARG,
                   "batched CG solve %d: grid (%d,%d,%d) out of range", k, S.nx, S.ny, S.nz);
        SW_REQUIRE(S.xchunk >= 1 && (S.xchunk % kCgSub == 0 || S.xchunk >= S.nx), SW_ERR_ARG,
                   "batched CG solve %d: xchunk %d must be a multiple of %d or at least nx", k, S.xchunk, kCgSub);
        SW_REQUIRE(S.f[0] && S.f[1] && S.f[2] && S.phi && S.r && S.p0 && S.p1 && S.part, SW_ERR_ARG,
                   "batched CG solve %d: null field or workspace", k);
        base[k + 1] = base[k] + (int64_t)((S.nx + kCgSub - 1) / kCgSub) * ((S.ny + kCgBY - 1) / kCgBY) *
                                    ((S.nz + 63) / 64);
    }
    std::vector<unsigned char> seen(base[n_solves], 0);
    for (int64_t j = 0; j < n_items; ++j) {
        const int4 t = it[j];
        SW_REQUIRE(t.x >= 0 && t.x < n_solves, SW_ERR_ARG, "CG item %lld: solve %d out of range", (long long)j, t.x);
        const sw_batch_cg& S = d[t.x];
        const int nty = (S.ny + kCgBY - 1) / kCgBY, ntz = (S.nz + 63) / 64;
        SW_REQUIRE(t.y >= 0 && t.y < ntz && t.z >= 0 && t.z < nty && t.w >= 0 && (int64_t)t.w * S.xchunk < S.nx,
                   SW_ERR_ARG, "CG item %lld: tile (%d,%d,%d) outside solve %d", (long long)j, t.y, t.z, t.w, t.x);
        const int x0 = t.w * S.xchunk, x1 = std::min(x0 + S.xchunk, S.nx);
        for (int x = x0; x < x1; x += kCgSub) {
            const int64_t slot = base[t.x] + ((int64_t)(x / kCgSub) * nty + t.z) * ntz + t.y;
            SW_REQUIRE(!seen[slot], SW_ERR_ARG, "CG item %lld: tile (%d,%d,%d) of solve %d listed twice",
                       (long long)j, t.y, t.z, t.w, t.x);
            seen[slot] = 1;
        }
    }
    for (int k = 0; k < n_solves; ++k)
        for (int64_t q = base[k]; q < base[k + 1]; ++q)
            SW_REQUIRE(seen[q], SW_ERR_ARG, "batched CG solve %d: its items do not cover the grid", k);
    return SW_OK;
}
}  // namespace

int sw_batch_project(const sw_batch_cg* solves, int32_t n_solves, const int32_t* items, int64_t n_items,
                     void* states, sw_cg_info* info, void* stream) {
    SW_REQUIRE(solves && items && info && states, SW_ERR_ARG, "null pointer");
    SW_REQUIRE(n_solves >= 1 && n_items >= 1, SW_ERR_ARG, "empty batch");
    hipStream_t s = (hipStream_t)stream;
    if (const int rc = check_batch_cg(solves, n_solves, items, n_items, s)) return rc;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_cg_dir_kernel, 64 * kCgBY, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const unsigned nb = (unsigned)(per_cu * device_cu_count());
    const dim3 blk(64, kCgBY);
    const int4* it = (const int4*)items;
    CgState* st = (CgState*)states;
    std::vector<CgState> h(n_solves);
    auto poll = [&]() -> int {
        SW_TRY_HIP(hipMemcpyAsync(h.data(), st, sizeof(CgState) * n_solves, hipMemcpyDeviceToHost, s));
        SW_TRY_HIP(hipStreamSynchronize(s));
        return SW_OK;
    };
    hipLaunchKernelGGL(batch_cg_reset_kernel, dim3((n_solves + 64) / 64), dim3(64), 0, s, solves, n_solves, st);
    hipLaunchKernelGGL(batch_cg_init_kernel, dim3(nb), blk, 0, s, solves, st, it, n_items);
    SW_LAUNCH_CHECK();
    int rc = poll();
    int64_t chunk = 8;
    while (rc == SW_OK) {
        // launches up to the next poll: no more than any running solve can still take
        int64_t left = 0;
        for (int k = 0; k < n_solves; ++k)
            if (h[k].stop == 0 && h[k].max_iters - h[k].iters > left) left = h[k].max_iters - h[k].iters;
        if (left <= 0) break;
        const int64_t n_launch = chunk < left ? chunk : left;
        for (int64_t j = 0; j < n_launch; ++j) {
            hipLaunchKernelGGL(batch_cg_dir_kernel, dim3(nb), blk, 0, s, solves, st, n_solves, it, n_items);
            hipLaunchKernelGGL(batch_cg_step_kernel, dim3(nb), blk, 0, s, solves, st, n_solves, it, n_items);
        }
        SW_LAUNCH_CHECK();
        rc = poll();
        if (chunk < 64) chunk *= 2;
    }
    if (rc != SW_OK) return rc;
    hipLaunchKernelGGL(batch_cg_flush_kernel, dim3(nb), blk, 0, s, solves, st, it, n_items);
    hipLaunchKernelGGL(batch_cg_grad_sub_kernel, dim3(nb), blk, 0, s, solves, it, n_items);
    SW_LAUNCH_CHECK();
    for (int k = 0; k < n_solves; ++k) {
        info[k].iters = h[k].iters;
        info[k].relres = h[k].relres;
        info[k].max_div = h[k].max_div;
        info[k].reason = h[k].reason;
        info[k].pad = 0;
        info[k].true_relres = h[k].true_relres;  // -1: the batched CG has no true-residual pass
    }
    return SW_OK;
}

}  // extern "C"
