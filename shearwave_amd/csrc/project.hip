// Helmholtz-Hodge projection of a force field on gfx950.
//
// Restates project_vector_field (solver.py:644-683) and poisson_cg
// (solver.py:1010-1050): mask the field, take its 2h central divergence,
// solve A phi = -div f with A = -div_c grad_c by conjugate gradients on the
// interior, subtract grad_c phi and mask again.
//
// The CG runs entirely on the device. Each iteration is two kernels:
//   dir:  p = r + beta p_old at the 13 points of the stride-2 stencil of A,
//         stores p, reduces p.Ap                      (reads r, p_old; writes p)
//   step: recomputes Ap from p, phi += alpha p, r -= alpha Ap, reduces r.r
//                                                     (reads phi, r, p; writes phi, r)
// i.e. 64 B per voxel per iteration in float64 instead of the textbook 88 B,
// with bit-identical arithmetic (Ap is recomputed by the same operations).
// With float32 vectors (r, p_old, p in float; phi in double) it is 40 B; the
// products, dot products and residual norm are float64 either way. The
// x-marching kernels defer phi to every other iteration (march_cg.h), which
// is bit-identical and takes one iteration pair from 2 x 64 to 112 B
// (float64 vectors) or from 2 x 40 to 68 B (float32).
// Scalars (alpha, beta, rs, stop flag) live in device memory; the last block
// of each kernel reduces the per-block partials in a fixed order (so results
// are deterministic) and applies the stopping rules. The host only polls the
// stop flag once per chunk of iterations.
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "march_cg.h"
#include "spectral.h"
#include "sw_common.h"

namespace sw {
namespace {

constexpr int kMaxRedBlocks = 2048;

__global__ void cg_reset_kernel(CgState* st, double tol, long long max_iters, double stage_tol, int precond) {
    cg_reset(st, tol, max_iters, stage_tol);
    st->precond = precond;
}

// r = -div_c(mask f) on the interior, p_old = 0, phi = 0; reduces r.r and max|div|.
template <typename T, typename V>
__global__ void __launch_bounds__(256) cg_init_kernel(const T* __restrict__ fx, const T* __restrict__ fy,
                                                      const T* __restrict__ fz, double* __restrict__ phi,
                                                      V* __restrict__ r, V* __restrict__ p_old, Grid g,
                                                      double i2hx, double i2hy, double i2hz, double* part_sum,
                                                      double* part_max, CgState* st) {
    __shared__ double red[16];
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    double s = 0.0, mx = 0.0;
    SW_FOR_EACH_VOXEL(g, {
        double rv = 0.0;
        if (g.interior(x, y, z)) {
            // neighbours on the boundary layer are masked to zero (solver.py:666-668)
            const double xp = (x + 1 < g.nx - 1) ? (double)fx[i + P] : 0.0;
            const double xm = (x - 1 > 0) ? (double)fx[i - P] : 0.0;
            const double yp = (y + 1 < g.ny - 1) ? (double)fy[i + R] : 0.0;
            const double ym = (y - 1 > 0) ? (double)fy[i - R] : 0.0;
            const double zp = (z + 1 < g.nz - 1) ? (double)fz[i + 1] : 0.0;
            const double zm = (z - 1 > 0) ? (double)fz[i - 1] : 0.0;
            const double div = (xp - xm) * i2hx + (yp - ym) * i2hy + (zp - zm) * i2hz;
            rv = as_stored<V>(-div);
            s += rv * rv;
            mx = fmax(mx, fabs(div));
        }
        r[i] = (V)rv;
        p_old[i] = (V)0;
        phi[i] = 0.0;
    });
    s = block_sum(s, red);
    mx = block_max(mx, red);
    if (publish_slot(part_sum, part_max, blockIdx.x, s, mx, &st->counter, gridDim.x)) {
        double tot, m;
        reduce_slots(part_sum, part_max, gridDim.x, red, &tot, &m);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_init(st, tot, m);
    }
}

// A x at an interior voxel for one axis: -(G(c+1) - G(c-1)) / (2h), G = 0 on the boundary layer.
// `v(k)` returns x at offset k*stride; c is the voxel's coordinate on the axis.
#define SW_AXIS_TERM(v, c, n, q) \
    ((((c) + 1 < (n) - 1) ? ((v)(2) - (v)(0)) : 0.0) - (((c) - 1 > 0) ? ((v)(0) - (v)(-2)) : 0.0)) * (q)

// True residual r = b - A phi in float64, with b = -div_c(mask f) formed as
// cg_init_kernel forms it but never rounded to V, and phi as accumulated (every
// deferred term added). Reduces r.r over the values as rounded to V. With
// `replace` the rounded values overwrite r and the last block restarts the
// recursion from them (cg_after_true_residual); otherwise r is left alone and
// only the norm is recorded.
template <typename T, typename V>
__global__ void __launch_bounds__(256) cg_true_residual_kernel(const T* __restrict__ fx, const T* __restrict__ fy,
                                                               const T* __restrict__ fz,
                                                               const double* __restrict__ phi, V* __restrict__ r,
                                                               Grid g, double i2hx, double i2hy, double i2hz,
                                                               double qx, double qy, double qz, double* part_sum,
                                                               CgState* st, int replace, double next_tol) {
    __shared__ double red[16];
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    double s = 0.0;
    SW_FOR_EACH_VOXEL(g, {
        if (g.interior(x, y, z)) {
            const double xp = (x + 1 < g.nx - 1) ? (double)fx[i + P] : 0.0;
            const double xm = (x - 1 > 0) ? (double)fx[i - P] : 0.0;
            const double yp = (y + 1 < g.ny - 1) ? (double)fy[i + R] : 0.0;
            const double ym = (y - 1 > 0) ? (double)fy[i - R] : 0.0;
            const double zp = (z + 1 < g.nz - 1) ? (double)fz[i + 1] : 0.0;
            const double zm = (z - 1 > 0) ? (double)fz[i - 1] : 0.0;
            const double b = -((xp - xm) * i2hx + (yp - ym) * i2hy + (zp - zm) * i2hz);
            const double fv = phi[i];
            auto px = [&](int k) { return k == 0 ? fv : phi[i + k * P]; };
            auto py = [&](int k) { return k == 0 ? fv : phi[i + k * R]; };
            auto pz = [&](int k) { return k == 0 ? fv : phi[i + k]; };
            const double aphi = -(SW_AXIS_TERM(px, x, g.nx, qx) + SW_AXIS_TERM(py, y, g.ny, qy) +
                                  SW_AXIS_TERM(pz, z, g.nz, qz));
            const V rv = (V)(b - aphi);
            s += (double)rv * (double)rv;
            if (replace) r[i] = rv;
        }
    });
    s = block_sum(s, red);
    if (publish_slot(part_sum, nullptr, blockIdx.x, s, 0.0, &st->counter, gridDim.x)) {
        double rs, unused;
        reduce_slots(part_sum, nullptr, gridDim.x, red, &rs, &unused);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_true_residual(st, rs, replace, next_tol);
    }
}

// Direction update fused with A p and p.Ap.
template <typename V>
__global__ void __launch_bounds__(256) cg_dir_kernel(const V* __restrict__ r, const V* __restrict__ p_old,
                                                     V* __restrict__ p, Grid g, double qx, double qy, double qz,
                                                     double* part_sum, CgState* st) {
    if (st->stop) return;
    __shared__ double red[16];
    const double beta = st->beta;
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    double s = 0.0;
    SW_FOR_EACH_VOXEL(g, {
        double pv = 0.0;
        if (g.interior(x, y, z)) {
            auto pn = [&](int64_t k) { return as_stored<V>(fma(beta, (double)p_old[i + k], (double)r[i + k])); };
            pv = pn(0);
            auto px = [&](int k) { return k == 0 ? pv : pn(k * P); };
            auto py = [&](int k) { return k == 0 ? pv : pn(k * R); };
            auto pz = [&](int k) { return k == 0 ? pv : pn((int64_t)k); };
            const double ap = -(SW_AXIS_TERM(px, x, g.nx, qx) + SW_AXIS_TERM(py, y, g.ny, qy) +
                                SW_AXIS_TERM(pz, z, g.nz, qz));
            s += pv * ap;
        }
        p[i] = (V)pv;
    });
    s = block_sum(s, red);
    if (publish_slot(part_sum, nullptr, blockIdx.x, s, 0.0, &st->counter, gridDim.x)) {
        double denom, unused;
        reduce_slots(part_sum, nullptr, gridDim.x, red, &denom, &unused);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_dir(st, denom);
    }
}

// phi += alpha p, r -= alpha A p (A p recomputed), r.r and the stopping rules.
template <typename V>
__global__ void __launch_bounds__(256) cg_step_kernel(double* __restrict__ phi, V* __restrict__ r,
                                                      const V* __restrict__ p, Grid g, double qx, double qy,
                                                      double qz, double* part_sum, CgState* st) {
    if (st->stop) return;
    __shared__ double red[16];
    const double alpha = st->alpha;
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    double s = 0.0;
    SW_FOR_EACH_VOXEL(g, {
        if (g.interior(x, y, z)) {
            const double pv = (double)p[i];
            auto px = [&](int k) { return k == 0 ? pv : (double)p[i + k * P]; };
            auto py = [&](int k) { return k == 0 ? pv : (double)p[i + k * R]; };
            auto pz = [&](int k) { return k == 0 ? pv : (double)p[i + k]; };
            const double ap = -(SW_AXIS_TERM(px, x, g.nx, qx) + SW_AXIS_TERM(py, y, g.ny, qy) +
                                SW_AXIS_TERM(pz, z, g.nz, qz));
            phi[i] = fma(alpha, pv, phi[i]);
            const V rv = (V)((double)r[i] - alpha * ap);
            r[i] = rv;
            s += (double)rv * (double)rv;
        }
    });
    s = block_sum(s, red);
    if (publish_slot(part_sum, nullptr, blockIdx.x, s, 0.0, &st->counter, gridDim.x)) {
        double rs_new, unused;
        reduce_slots(part_sum, nullptr, gridDim.x, red, &rs_new, &unused);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_step(st, rs_new);
    }
}

// Single-solve x-marching CG kernels: one tile per block (tiles z fastest,
// then y, then x chunk, XCD-aware), one partial per block.
__device__ inline CgTile cg_block_tile(const Grid& g, int xchunk, int nty, int ntz) {
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);  // XCD-aware: neighbours share an L2
    const int tzi = (int)(b % (unsigned)ntz);                   // z fastest: whole z rows per XCD
    const unsigned rest = b / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int txi = (int)(rest / (unsigned)nty);
    const int x0 = txi * xchunk;
    return cg_tile(g, x0, min(x0 + xchunk, g.nx), tyi * kCgBY, tzi * 64);
}

template <typename V>
__global__ void __launch_bounds__(64 * kCgBY) cg_dir_march_kernel(const V* __restrict__ r, const V* __restrict__ p_old,
                                                                  V* __restrict__ p, Grid g, double qx, double qy,
                                                                  double qz, int xchunk, int nty, int ntz,
                                                                  double* part_sum, CgState* st) {
    if (st->stop) return;
    __shared__ CgLdsT<V> lds;
    __shared__ double red[16];
    const CgTile t = cg_block_tile(g, xchunk, nty, ntz);
    double s = cg_dir_tile<V>(r, p_old, p, g, st->beta, qx, qy, qz, t, lds);
    s = block_sum(s, red);
    if (publish_slot(part_sum, nullptr, blockIdx.x, s, 0.0, &st->counter, gridDim.x)) {
        double denom, unused;
        reduce_slots(part_sum, nullptr, gridDim.x, red, &denom, &unused);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_dir(st, denom);
    }
}

// PHI: kPhiSkip on odd iterations, kPhiPair (p_prev = p of the odd one) on even ones.
template <typename V, int PHI>
__global__ void __launch_bounds__(64 * kCgBY) cg_step_march_kernel(double* __restrict__ phi, V* __restrict__ r,
                                                                   const V* __restrict__ p,
                                                                   const V* __restrict__ p_prev, Grid g, double qx,
                                                                   double qy, double qz, int xchunk, int nty, int ntz,
                                                                   double* part_sum, CgState* st) {
    if (st->stop) return;
    __shared__ CgLdsT<V> lds;
    __shared__ double red[16];
    const CgTile t = cg_block_tile(g, xchunk, nty, ntz);
    double s = cg_step_tile<V, PHI>(phi, r, p, p_prev, g, st->alpha, st->alpha_prev, qx, qy, qz, t, lds);
    s = block_sum(s, red);
    if (publish_slot(part_sum, nullptr, blockIdx.x, s, 0.0, &st->counter, gridDim.x)) {
        double rs_new, unused;
        reduce_slots(part_sum, nullptr, gridDim.x, red, &rs_new, &unused);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_step(st, rs_new);
    }
}

// After the march CG: the deferred phi term of a solve whose last iteration
// left its term pending (an odd number of iterations since phi was last
// complete; the host launches it only then). p_last is that iteration's p.
template <typename V>
__global__ void __launch_bounds__(64 * kCgBY) cg_flush_march_kernel(double* __restrict__ phi,
                                                                    const V* __restrict__ p_last, Grid g, int xchunk,
                                                                    int nty, int ntz, const CgState* st) {
    const CgTile t = cg_block_tile(g, xchunk, nty, ntz);
    cg_flush_phi_tile<V>(phi, p_last, g, st->alpha, t);
}

// One 1024-thread block per CU: long x chunks keep the per-chunk prologue
// (six planes of p, r, p_old loaded before the first output) small.
#ifndef SW_CG_BLOCKS_PER_CU
#define SW_CG_BLOCKS_PER_CU 1
#endif

// Shortest x chunk of the march, in planes: a chunk re-reads the planes of
// its prologue, so short chunks cost traffic on large grids.
#ifndef SW_CG_MIN_XC
#define SW_CG_MIN_XC 12
#endif
constexpr int kCgMinXc = SW_CG_MIN_XC;

struct CgShape {
    int xchunk, nty, ntz, ntx;
    unsigned blocks() const { return (unsigned)nty * ntz * ntx; }
};

CgShape cg_shape(const Grid& g) {
    CgShape c;
    c.nty = (g.ny + kCgBY - 1) / kCgBY;
    c.ntz = (g.nz + 63) / 64;
    const int64_t cols = (int64_t)c.nty * c.ntz;
    // about SW_CG_BLOCKS_PER_CU blocks per CU in one wave, chunks of at least kCgMinXc planes
    const int64_t target = (int64_t)device_cu_count() * SW_CG_BLOCKS_PER_CU;
    int64_t ntx = (target + cols - 1) / cols;
    if (ntx < 1) ntx = 1;
    int64_t xc = (g.nx + ntx - 1) / ntx;
    if (xc < kCgMinXc) xc = kCgMinXc;
    c.xchunk = (int)xc;
    c.ntx = (g.nx + c.xchunk - 1) / c.xchunk;
    return c;
}

// Upper bound on the march grid's block count (chunks never shorter than kCgMinXc planes).
int64_t cg_max_blocks(int64_t nx, int64_t ny, int64_t nz) {
    return ((ny + kCgBY - 1) / kCgBY) * ((nz + 63) / 64) * ((nx + kCgMinXc - 1) / kCgMinXc);
}

// f_S = mask (mask f - grad_c phi); optionally out = scale * f_S * inv_rho.
template <typename T>
__global__ void __launch_bounds__(256) grad_sub_kernel(T* __restrict__ fx, T* __restrict__ fy, T* __restrict__ fz,
                                                       const double* __restrict__ phi, Grid g, double i2hx,
                                                       double i2hy, double i2hz, T* __restrict__ ox,
                                                       T* __restrict__ oy, T* __restrict__ oz,
                                                       const T* __restrict__ inv_rho, double scale) {
    const int64_t P = g.plane();
    const int64_t R = g.nz;
    SW_FOR_EACH_VOXEL(g, {
        double a = 0.0, b = 0.0, c = 0.0;
        if (g.interior(x, y, z)) {
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
    });
}

// phi is float64; r, p0 and p1 hold float64 or float32 vectors (each slot is
// sized for float64, so one workspace size serves both).
struct Workspace {
    double* phi;
    void *r, *p0, *p1;
    double *part_sum, *part_max;
    CgState* st;
    float* tables;  // spectral preconditioner tables (when the workspace holds them), else null
};

int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

int64_t partial_capacity(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t m = cg_max_blocks(nx, ny, nz);
    return m > kMaxRedBlocks ? m : kMaxRedBlocks;
}

// Without the preconditioner's tables.
int64_t plain_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t n = nx * ny * nz;
    return 4 * align256(n * 8) + 2 * align256(partial_capacity(nx, ny, nz) * 8) + 256;
}

// With them (grids whose chains have tables).
int64_t workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t b = plain_workspace_bytes(nx, ny, nz);
    return spectral_shape(nx, ny, nz) ? b + align256(spectral_table_bytes(nx, ny, nz)) : b;
}

Workspace carve(void* base, int64_t nx, int64_t ny, int64_t nz) {
    const int64_t n = nx * ny * nz;
    const int64_t pc = partial_capacity(nx, ny, nz);
    char* p = (char*)base;
    Workspace w;
    w.phi = (double*)p;
    p += align256(n * 8);
    w.r = p;
    p += align256(n * 8);
    w.p0 = p;
    p += align256(n * 8);
    w.p1 = p;
    p += align256(n * 8);
    w.part_sum = (double*)p;
    p += align256(pc * 8);
    w.part_max = (double*)p;
    p += align256(pc * 8);
    w.st = (CgState*)p;
    p += 256;
    w.tables = spectral_shape(nx, ny, nz) ? (float*)p : nullptr;
    return w;
}

unsigned red_grid(const Grid& g) {
    int64_t cap = (int64_t)device_cu_count() * 8;
    if (cap > kMaxRedBlocks) cap = kMaxRedBlocks;
    return row_grid(g, cap);
}

// The preconditioned CG is used only at tolerances this tight or tighter.
constexpr double kPcgMaxTol = 1e-7;

// Iterations the plain CG takes to tol on a grid, from measurements (256^3 at
// 1e-8: 599; 48^3 at 1e-8: 113; 40x21x70 at 1e-9: 179), rounded up: the
// preconditioned CG stands in for it only where max_iters would not stop it.
inline int64_t plain_cg_iters(const Grid& g, double tol) {
    if (!(tol > 0.0)) return INT64_MAX;
    const int nmax = g.nx > g.ny ? (g.nx > g.nz ? g.nx : g.nz) : (g.ny > g.nz ? g.ny : g.nz);
    return (int64_t)ceil(nmax * (0.25 * log10(1.0 / tol) + 0.5));
}

// Whether a projection (with a workspace of workspace_bytes) takes the spectral
// preconditioner: float32 vectors, a grid with tables, and tol / max_iters at
// which the plain CG converges too; SW_CG_PRECOND=0 / 1 turns it off / on
// wherever possible.
inline bool uses_preconditioner(int vec_dtype, int64_t nx, int64_t ny, int64_t nz, double tol, int64_t max_iters) {
    if (vec_dtype != SW_F32 || !spectral_shape(nx, ny, nz)) return false;
    const char* env = getenv("SW_CG_PRECOND");
    if (env && env[0] == '0') return false;
    if (env && env[0] == '1') return true;
    return tol <= kPcgMaxTol && max_iters >= plain_cg_iters(Grid{(int)nx, (int)ny, (int)nz}, tol);
}

// Stop levels of the float32-vector CG before its true-residual checks: each
// level is this factor below the one before, down to the caller's tol. One
// replacement per decade keeps each replacement's change to r (the drift since
// the last one, ~eps32 * 10 relative) too small to disturb the directions'
// conjugacy: on a 40x21x70 test field the iteration counts then equal float64
// vectors' down to tol 1e-9 (169 at 1e-8, 179 at 1e-9), where one replacement
// per four decades needed 169 and 199.
constexpr double kStageFactor = 0.1;
// Residual replacements after which the solve ends with whatever it reached.
constexpr int kMaxReplacements = 24;

// Launch the CG from r0 to its stop on the stream; returns the final state
// in hs (host). V is the storage of r and p.
//
// Float64 vectors iterate to tol, as poisson_cg does. With float32 vectors the
// recursively updated r drifts from b - A phi by float32 rounding (r0 itself
// is b rounded), so the recursion stops first at a looser stage level; there
// the true residual b - A phi is formed in float64 and replaces r (residual
// replacement), and iteration continues to the next level, down to tol, until
// the TRUE residual meets tol. Each replacement removes the drift accumulated
// so far, which scales with the residual at that point, so the final drift is
// ~ eps32 * (last stage level) relative to r0, far below tol. Either way a
// last pass records the true residual of the returned phi.
template <typename T, typename V>
int run_cg(const Workspace& w, const Grid& g, bool march, const CgShape& cs, unsigned nb, const T* fx, const T* fy,
           const T* fz, double i2hx, double i2hy, double i2hz, double qx, double qy, double qz, double tol,
           int64_t max_iters, double stage, const SubLat* pcg, CgState& hs, hipStream_t s) {
    const dim3 blk = row_block();
    const bool replacing = sizeof(V) == 4;
    V* const p_odd = (V*)w.p1;   // iteration k (counted from 1) writes its p here when k is odd
    V* const p_even = (V*)w.p0;  // and here when k is even
    V* r = (V*)w.r;
    // preconditioned: z = M r in the upper half of r's slot, the transforms' work in those of p0 and p1
    V* const z = pcg ? r + g.size() : r;
    float* const W1 = (float*)w.p0 + g.size();
    float* const W2 = (float*)w.p1 + g.size();
    auto precondition = [&]() {
        if (pcg) spectral_apply<V>(r, z, W1, W2, g, *pcg, nb, w.part_sum, w.st, s);
    };
    precondition();  // z0 = M r0 and r0.z0 (p1 = z0: p_old starts at zero)
    SW_LAUNCH_CHECK();
    SW_TRY_HIP(hipMemcpyAsync(&hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    SW_TRY_HIP(hipStreamSynchronize(s));
    int64_t base = 0;  // iterations whose phi terms are all in phi (the pairing restarts after a flush)
    int64_t chunk = pcg ? 2 : 8;  // iterations launched between two polls of the stop flag
    for (int round = 0;; ++round) {
        int64_t launched = hs.iters;
        while (!hs.stop && launched < max_iters) {
            const int64_t k = (max_iters - launched < chunk) ? (max_iters - launched) : chunk;
            for (int64_t j = 0; j < k; ++j) {
                const int64_t it = launched + j + 1;
                V* p_new = (it & 1) ? p_odd : p_even;
                V* p_old = (it & 1) ? p_even : p_odd;
                if (march) {
                    const bool skip = ((it - base) & 1) == 1;
                    hipLaunchKernelGGL(cg_dir_march_kernel<V>, dim3(cs.blocks()), dim3(64, kCgBY), 0, s, z, p_old,
                                       p_new, g, qx, qy, qz, cs.xchunk, cs.nty, cs.ntz, w.part_sum, w.st);
                    if (skip)
                        hipLaunchKernelGGL((cg_step_march_kernel<V, kPhiSkip>), dim3(cs.blocks()), dim3(64, kCgBY), 0,
                                           s, w.phi, r, p_new, p_old, g, qx, qy, qz, cs.xchunk, cs.nty, cs.ntz,
                                           w.part_sum, w.st);
                    else
                        hipLaunchKernelGGL((cg_step_march_kernel<V, kPhiPair>), dim3(cs.blocks()), dim3(64, kCgBY), 0,
                                           s, w.phi, r, p_new, p_old, g, qx, qy, qz, cs.xchunk, cs.nty, cs.ntz,
                                           w.part_sum, w.st);
                } else {
                    hipLaunchKernelGGL(cg_dir_kernel<V>, dim3(nb), blk, 0, s, z, p_old, p_new, g, qx, qy, qz,
                                       w.part_sum, w.st);
                    hipLaunchKernelGGL(cg_step_kernel<V>, dim3(nb), blk, 0, s, w.phi, r, p_new, g, qx, qy, qz,
                                       w.part_sum, w.st);
                }
                precondition();
            }
            SW_LAUNCH_CHECK();
            launched += k;
            SW_TRY_HIP(hipMemcpyAsync(&hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
            SW_TRY_HIP(hipStreamSynchronize(s));
            if (chunk < 64) chunk *= 2;
        }
        // hs is final here: kernels launched after the stop returned at once
        if (march && ((hs.iters - base) & 1)) {
            hipLaunchKernelGGL(cg_flush_march_kernel<V>, dim3(cs.blocks()), dim3(64, kCgBY), 0, s, w.phi,
                               (const V*)((hs.iters & 1) ? p_odd : p_even), g, cs.xchunk, cs.nty, cs.ntz, w.st);
            SW_LAUNCH_CHECK();
        }
        base = hs.iters;
        if (hs.reason == SW_CG_EARLY_EXIT) break;  // phi = 0 solves the system exactly
        const bool replace = replacing && hs.reason == SW_CG_CONVERGED && round < kMaxReplacements;
        double next = stage * kStageFactor;
        if (next < tol) next = tol;
        hipLaunchKernelGGL((cg_true_residual_kernel<T, V>), dim3(nb), blk, 0, s, fx, fy, fz, (const double*)w.phi, r,
                           g, i2hx, i2hy, i2hz, qx, qy, qz, w.part_sum, w.st, replace ? 1 : 0, next);
        SW_LAUNCH_CHECK();
        SW_TRY_HIP(hipMemcpyAsync(&hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        SW_TRY_HIP(hipStreamSynchronize(s));
        if (!replace || hs.stop) break;
        stage = next;
        precondition();  // z = M r of the replaced residual; beta from r.z
        SW_LAUNCH_CHECK();
    }
    return SW_OK;
}

}  // namespace
}  // namespace sw

// ------------------------------------------------- no-allocation projection
//
// Scratch of sw_project_vector_field_noalloc: the CG workspace (phi, r, p,
// p_old, the partial slots and the CG state) for grids whose workspace fits,
// i.e. up to 2^20 voxels. Each GPU the library runs on holds its own copy
// until the process ends. Calls on one GPU share it, so calls that overlap on
// one GPU (different streams or threads) race. Nothing in it needs to start at
// zero: every call resets the CG state and writes phi, r and p_old before any
// read, and each slot is written before it is summed.
constexpr int64_t kNoAllocScratchBytes = (int64_t(33) << 20) + (int64_t(320) << 10);  // incl. spectral tables
constexpr int64_t kNoAllocMaxVoxels = int64_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace sw {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device ordinal (null until looked up)

// The current device's address of `scratch`, looked up once per device. The
// array's host-side name is not a device address and is never passed on.
int noalloc_scratch(void** out) {
    int dev = -1;
    SW_TRY_HIP(hipGetDevice(&dev));
    SW_REQUIRE(dev >= 0 && dev < kMaxDevices, SW_ERR_ARG, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SW_TRY_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
        SW_REQUIRE(p && ((uintptr_t)p & 255) == 0, SW_ERR_HIP, "scratch has no 256-byte aligned device address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SW_OK;
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int64_t sw_project_workspace_bytes(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    return workspace_bytes(nx, ny, nz);
}

int sw_project_vector_field(int dtype, int vec_dtype, void* fx, void* fy, void* fz, int64_t nx, int64_t ny,
                            int64_t nz, double dx, double dy, double dz, double tol, int64_t max_iters,
                            void* workspace, int64_t workspace_bytes_given, void* const* scale_out,
                            const void* inv_rho, double scale, sw_cg_info* info, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_dtype(vec_dtype))) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    SW_REQUIRE(fx && fy && fz, SW_ERR_ARG, "null force component");
    SW_REQUIRE(max_iters >= 0, SW_ERR_ARG, "max_iters must be >= 0");
    SW_REQUIRE(dx > 0 && dy > 0 && dz > 0, SW_ERR_ARG, "grid spacing must be positive");
    SW_REQUIRE(workspace && workspace_bytes_given >= plain_workspace_bytes(nx, ny, nz), SW_ERR_WORKSPACE,
               "workspace of %lld bytes is smaller than the %lld needed", (long long)workspace_bytes_given,
               (long long)workspace_bytes(nx, ny, nz));
    SW_REQUIRE(((uintptr_t)workspace & 255) == 0, SW_ERR_ARG, "workspace must be 256-byte aligned");
    if (scale_out) SW_REQUIRE(scale_out[0] && scale_out[1] && scale_out[2], SW_ERR_ARG, "null scale_out component");

    hipStream_t s = (hipStream_t)stream;
    Grid g{(int)nx, (int)ny, (int)nz};
    Workspace w = carve(workspace, nx, ny, nz);
    const unsigned nb = red_grid(g);
    const CgShape cs = cg_shape(g);
    // the march kernels need at least one 64-wide z tile and 8 rows; tiny grids use the direct kernels
    const char* env_k = getenv("SW_CG_KERNEL");
    bool march = nz >= 32 && ny >= 8 && g.plane() < kMaxMarchPlane;
    if (env_k && env_k[0] == 'd') march = false;
    const dim3 blk = row_block();
    const double i2hx = 1.0 / (2.0 * dx), i2hy = 1.0 / (2.0 * dy), i2hz = 1.0 / (2.0 * dz);
    const double qx = 1.0 / (4.0 * dx * dx), qy = 1.0 / (4.0 * dy * dy), qz = 1.0 / (4.0 * dz * dz);

    // the spectral preconditioner (float32 vectors): where the workspace holds its tables and the plain CG
    // would meet tol too (tol <= 1e-7, max_iters at least its iteration estimate), so that the result is the
    // reference's to well within the float32 form's tolerance; SW_CG_PRECOND=0 / 1 turns it off / on
    const bool pcg = workspace_bytes_given >= workspace_bytes(nx, ny, nz) &&
                     uses_preconditioner(vec_dtype, nx, ny, nz, tol, max_iters);
    if (!pcg) w.tables = nullptr;
    SubLat S{};
    if (pcg && (rc = spectral_setup(nx, ny, nz, dx, dy, dz, w.tables, S, s))) return rc;
    // float32 vectors first stop at a looser stage level for a residual replacement (run_cg); the
    // preconditioned CG converges in a few iterations and checks its true residual at tol
    const double stage = (vec_dtype == SW_F32 && tol < kStageFactor && !pcg) ? kStageFactor : tol;
    hipLaunchKernelGGL(cg_reset_kernel, dim3(1), dim3(1), 0, s, w.st, tol, (long long)max_iters, stage, pcg ? 1 : 0);
    SW_LAUNCH_CHECK();
    CgState hs;
    if (vec_dtype == SW_F32) {
        SW_DISPATCH(dtype, T, {
            hipLaunchKernelGGL((cg_init_kernel<T, float>), dim3(nb), blk, 0, s, (const T*)fx, (const T*)fy,
                               (const T*)fz, w.phi, (float*)w.r, (float*)w.p0, g, i2hx, i2hy, i2hz, w.part_sum,
                               w.part_max, w.st);
            SW_LAUNCH_CHECK();
            if ((rc = run_cg<T, float>(w, g, march, cs, nb, (const T*)fx, (const T*)fy, (const T*)fz, i2hx, i2hy,
                                       i2hz, qx, qy, qz, tol, max_iters, stage, pcg ? &S : nullptr, hs, s)))
                return rc;
        });
    } else {
        SW_DISPATCH(dtype, T, {
            hipLaunchKernelGGL((cg_init_kernel<T, double>), dim3(nb), blk, 0, s, (const T*)fx, (const T*)fy,
                               (const T*)fz, w.phi, (double*)w.r, (double*)w.p0, g, i2hx, i2hy, i2hz, w.part_sum,
                               w.part_max, w.st);
            SW_LAUNCH_CHECK();
            if ((rc = run_cg<T, double>(w, g, march, cs, nb, (const T*)fx, (const T*)fy, (const T*)fz, i2hx, i2hy,
                                        i2hz, qx, qy, qz, tol, max_iters, stage, nullptr, hs, s)))
                return rc;
        });
    }

    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(grad_sub_kernel<T>, dim3(row_grid(g, 0)), blk, 0, s, (T*)fx, (T*)fy, (T*)fz, w.phi,
                                   g, i2hx, i2hy, i2hz, scale_out ? (T*)scale_out[0] : nullptr,
                                   scale_out ? (T*)scale_out[1] : nullptr, scale_out ? (T*)scale_out[2] : nullptr,
                                   (const T*)inv_rho, scale));
    SW_LAUNCH_CHECK();
    if (info) {
        info->iters = hs.iters;
        info->relres = hs.relres;
        info->max_div = hs.max_div;
        info->reason = hs.stop ? hs.reason : SW_CG_MAX_ITERS;
        info->pad = 0;
        info->true_relres = hs.reason == SW_CG_EARLY_EXIT ? 0.0 : hs.true_relres;
    }
    return SW_OK;
}

int sw_project_preconditioned(int vec_dtype, int64_t nx, int64_t ny, int64_t nz, double tol, int64_t max_iters) {
    if (check_grid(nx, ny, nz)) return 0;
    return uses_preconditioner(vec_dtype, nx, ny, nz, tol, max_iters) ? 1 : 0;
}

int64_t sw_project_noalloc_max_voxels(void) { return kNoAllocMaxVoxels; }

int sw_project_vector_field_noalloc(int dtype, int vec_dtype, void* fx, void* fy, void* fz, int64_t nx, int64_t ny,
                                    int64_t nz, double dx, double dy, double dz, double tol, int64_t max_iters,
                                    void* const* scale_out, const void* inv_rho, double scale, sw_cg_info* info,
                                    void* stream) {
    int rc = check_grid(nx, ny, nz);
    if (rc) return rc;
    const int64_t n = nx * ny * nz;
    SW_REQUIRE(n <= kNoAllocMaxVoxels && workspace_bytes(nx, ny, nz) <= kNoAllocScratchBytes, SW_ERR_ARG,
               "grid %lldx%lldx%lld (%lld voxels) exceeds the no-allocation projection's scratch (%lld voxels, "
               "%lld bytes); use sw_project_vector_field with a workspace",
               (long long)nx, (long long)ny, (long long)nz, (long long)n, (long long)kNoAllocMaxVoxels,
               (long long)kNoAllocScratchBytes);
    void* ws = nullptr;
    if ((rc = noalloc_scratch(&ws))) return rc;
    return sw_project_vector_field(dtype, vec_dtype, fx, fy, fz, nx, ny, nz, dx, dy, dz, tol, max_iters, ws,
                                   kNoAllocScratchBytes, scale_out, inv_rho, scale, info, stream);
}

}  // extern "C"
