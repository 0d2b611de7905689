// Spectral preconditioner of the projection's CG on gfx950.
//
// The operator A = -div_c grad_c of poisson_cg (solver.py:1010-1050, with the
// masking of project_vector_field, solver.py:664-682) only couples points two
// apart along each axis, and on the interior it is the Kronecker sum of three
// 1-D operators, one per axis. Along an axis of n points the interior 1..n-2
// splits into two chains (odd points 1, 3, ..; even points 2, 4, ..); on a
// chain of m points the 1-D operator is (1/(4h^2)) T, T = tridiag(-1, 2, -1),
// except that a chain ending next to a masked flux (point 1, point n-2) has 1
// on the diagonal there (the flux G = grad_c phi is 0 on the boundary layer)
// where a chain ending next to a masked value (point 2, point n-3) keeps 2 (phi
// is 0 on the boundary layer). The interior is thus 8 independent
// sub-lattices (one chain per axis), and T has closed-form orthonormal
// eigenvectors for each pair of end types:
//   flux-flux    cos(k pi (t + 1/2) / m),             lambda = 2 - 2 cos(k pi / m)
//   flux-value   cos((2k+1) pi (t + 1/2) / (2m + 1)), lambda = 2 - 2 cos((2k+1) pi / (2m + 1))
//   value-flux   sin((2k+1) pi (t + 1) / (2m + 1)),   lambda as flux-value
//   value-value  sin((k+1) pi (t + 1) / (m + 1)),     lambda = 2 - 2 cos((k+1) pi / (m + 1))
// (t the point, k the mode; columns normalised numerically). So A^+ r on a
// sub-lattice is Q (Q^T r / Lambda) with Q = Qx (x) Qy (x) Qz and Lambda the
// sum of the three axes' lambda / (4 h^2) (the pseudo-inverse drops the zero
// mode of a sub-lattice whose three chains are flux-flux).
//
// The preconditioner z = A^+ r (float32) is one gather of r into the packed
// sub-lattices, three forward axis transforms (small GEMMs against the bases,
// the last one dividing by Lambda), three inverse ones and a scatter that
// also reduces r.z. All kernels return at once once the CG has stopped.
#pragma once

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "march_cg.h"
#include "sw_common.h"

namespace sw {

// Longest chain the tables are built for (axes of up to 2 * 256 + 3 points).
constexpr int kSpecMaxChain = 256;

// Chain c (0: odd points, 1: even points) of an axis of n points.
struct Chain {
    int m;          // points
    bool flux_lo;   // first point is 1 (its lower flux is masked)
    bool flux_hi;   // last point is n-2 (its upper flux is masked)
};

inline Chain chain_of(int n, int c) {
    Chain ch;
    const int first = 1 + c, last_max = n - 2;
    ch.m = last_max >= first ? (last_max - first) / 2 + 1 : 0;
    ch.flux_lo = c == 0;
    ch.flux_hi = ch.m > 0 && first + 2 * (ch.m - 1) == last_max;
    return ch;
}

// The packed sub-lattices of a grid and their tables (device pointers).
struct SubLat {
    int m[3][2];              // chain lengths per axis and parity
    long long off[8];         // offset of sub-lattice s = 4 cx + 2 cy + cz in the work buffers
    const float* q[3][2];     // bases: q[t * m + k] is mode k at point t
    const float* lam[3][2];   // eigenvalues of T
    float scale[3];           // 1 / (4 h^2) per axis
};

// Whether the spectral tables exist for a grid (every chain at most kSpecMaxChain points).
inline bool spectral_shape(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 2; ++c)
            if (chain_of((int)n[a], c).m > kSpecMaxChain) return false;
    return nx >= 3 && ny >= 3 && nz >= 3;
}

// Bytes of the tables of a grid (floats: m^2 + m per axis and chain).
inline int64_t spectral_table_bytes(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t n[3] = {nx, ny, nz};
    int64_t f = 0;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 2; ++c) {
            const int64_t m = chain_of((int)n[a], c).m;
            f += m * m + m;
        }
    return f * 4;
}

// Host tables of one chain: basis (row-major, t * m + k) and eigenvalues, built
// in double and rounded once; cached per (m, end types).
inline const std::vector<float>& chain_table(const Chain& ch) {
    static std::mutex mu;
    static std::map<int, std::vector<float>> cache;
    const int key = ch.m * 4 + (ch.flux_lo ? 2 : 0) + (ch.flux_hi ? 1 : 0);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const int m = ch.m;
    std::vector<double> q((size_t)m * m), lam(m);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < m; ++k) {
        double th;
        if (ch.flux_lo && ch.flux_hi)
            th = pi * k / m;
        else if (ch.flux_lo != ch.flux_hi)
            th = pi * (2 * k + 1) / (2.0 * m + 1.0);
        else
            th = pi * (k + 1) / (m + 1.0);
        lam[k] = 2.0 - 2.0 * cos(th);
        double nrm = 0.0;
        for (int t = 0; t < m; ++t) {
            const double v = ch.flux_lo ? cos(th * (t + 0.5)) : sin(th * (t + 1.0));
            q[(size_t)t * m + k] = v;
            nrm += v * v;
        }
        nrm = 1.0 / sqrt(nrm);
        for (int t = 0; t < m; ++t) q[(size_t)t * m + k] *= nrm;
    }
    std::vector<float> out((size_t)m * m + m);
    for (size_t j = 0; j < q.size(); ++j) out[j] = (float)q[j];
    for (int k = 0; k < m; ++k) out[(size_t)m * m + k] = (float)lam[k];
    return cache.emplace(key, std::move(out)).first->second;
}

// Lay the tables out at `tables` (device), upload them on stream s, and fill S.
inline int spectral_setup(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz, float* tables,
                          SubLat& S, hipStream_t s) {
    const int64_t n[3] = {nx, ny, nz};
    const double h[3] = {dx, dy, dz};
    float* p = tables;
    for (int a = 0; a < 3; ++a) {
        S.scale[a] = (float)(1.0 / (4.0 * h[a] * h[a]));
        for (int c = 0; c < 2; ++c) {
            const Chain ch = chain_of((int)n[a], c);
            S.m[a][c] = ch.m;
            S.q[a][c] = p;
            S.lam[a][c] = p + (size_t)ch.m * ch.m;
            if (ch.m > 0) {
                const std::vector<float>& t = chain_table(ch);
                SW_TRY_HIP(hipMemcpyAsync(p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, s));
            }
            p += (size_t)ch.m * ch.m + ch.m;
        }
    }
    long long off = 0;
    for (int sl = 0; sl < 8; ++sl) {
        S.off[sl] = off;
        off += (long long)S.m[0][sl >> 2] * S.m[1][(sl >> 1) & 1] * S.m[2][sl & 1];
    }
    return SW_OK;
}

namespace {

// r (interior voxels) into the packed sub-lattices W.
template <typename V>
__global__ void __launch_bounds__(256) spec_gather_kernel(const V* __restrict__ r, float* __restrict__ W, Grid g,
                                                          SubLat S, const CgState* st) {
    if (st->stop) return;
    SW_FOR_EACH_VOXEL(g, {
        if (g.interior(x, y, z)) {
            const int cx = (x - 1) & 1, cy = (y - 1) & 1, cz = (z - 1) & 1;
            const int my = S.m[1][cy], mz = S.m[2][cz];
            const int64_t j = S.off[4 * cx + 2 * cy + cz] + ((int64_t)((x - 1) >> 1) * my + ((y - 1) >> 1)) * mz +
                              ((z - 1) >> 1);
            W[j] = (float)r[i];
        }
    });
}

constexpr int kSpecTK = 16;  // K chunk of the transform tiles (64 x 64 outputs, 256 threads, 4 x 4 each)

// Transform along x (AXIS 0) or y (AXIS 1) of every sub-lattice (blockIdx.z):
// C_b[j][l] = sum_i A[j][i] X_b[i][l], with A = Q^T (forward) or Q (INV), X_b
// the (m x L) matrix of lines along the axis (x: one, L = my mz; y: one per x
// point, L = mz). blockIdx.y = (b, j tile), blockIdx.x = l tile.
template <int AXIS, bool INV>
__global__ void __launch_bounds__(256) spec_rows_kernel(const float* __restrict__ in, float* __restrict__ out, SubLat S,
                                                        const CgState* st) {
    if (st->stop) return;
    const int sl = blockIdx.z;
    const int cx = sl >> 2, cy = (sl >> 1) & 1, cz = sl & 1;
    const int mx = S.m[0][cx], my = S.m[1][cy], mz = S.m[2][cz];
    const int m = AXIS == 0 ? mx : my;
    const int L = AXIS == 0 ? my * mz : mz;
    const int nb = AXIS == 0 ? 1 : mx;
    if (m == 0 || L == 0) return;
    const int ntj = (m + 63) / 64;
    const int b = (int)(blockIdx.y / ntj), j0 = (int)(blockIdx.y % ntj) * 64, l0 = (int)blockIdx.x * 64;
    if (b >= nb || l0 >= L) return;
    const float* __restrict__ Q = S.q[AXIS][AXIS == 0 ? cx : cy];
    const int64_t base = S.off[sl] + (int64_t)b * m * L;
    const float* __restrict__ X = in + base;
    float* __restrict__ C = out + base;
    __shared__ float As[kSpecTK][64 + 4];
    __shared__ float Xs[kSpecTK][64 + 4];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    float acc[4][4] = {};
    for (int i0 = 0; i0 < m; i0 += kSpecTK) {
        for (int e = tid; e < kSpecTK * 64; e += 256) {
            const int ii = e >> 6, q = e & 63, i = i0 + ii, j = j0 + q, l = l0 + q;
            As[ii][q] = (i < m && j < m) ? (INV ? Q[(int64_t)j * m + i] : Q[(int64_t)i * m + j]) : 0.f;
            Xs[ii][q] = (i < m && l < L) ? X[(int64_t)i * L + l] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int ii = 0; ii < kSpecTK; ++ii) {
            const float4 a = *reinterpret_cast<const float4*>(&As[ii][ty * 4]);
            const float4 xv = *reinterpret_cast<const float4*>(&Xs[ii][tx * 4]);
            const float av[4] = {a.x, a.y, a.z, a.w}, xs[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fmaf(av[r], xs[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = j0 + ty * 4 + r;
        if (j >= m) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int l = l0 + tx * 4 + c;
            if (l < L) C[(int64_t)j * L + l] = acc[r][c];
        }
    }
}

// Transform along z (contiguous lines) of every sub-lattice: C[l][k] = sum_i
// X[l][i] B[i][k], with B = Q (forward) or Q^T (INV), X the (mx my x mz)
// matrix of lines. With SCALE (the last forward transform) the result is
// divided by Lambda = sum over axes of lambda / (4 h^2) (0 for a zero mode).
template <bool INV, bool SCALE>
__global__ void __launch_bounds__(256) spec_cols_kernel(const float* __restrict__ in, float* __restrict__ out, SubLat S,
                                                        const CgState* st) {
    if (st->stop) return;
    const int sl = blockIdx.z;
    const int cx = sl >> 2, cy = (sl >> 1) & 1, cz = sl & 1;
    const int mx = S.m[0][cx], my = S.m[1][cy], m = S.m[2][cz];
    const int L = mx * my;
    if (m == 0 || L == 0) return;
    const int k0 = (int)blockIdx.x * 64, l0 = (int)blockIdx.y * 64;
    if (k0 >= m || l0 >= L) return;
    const float* __restrict__ Q = S.q[2][cz];
    const float* __restrict__ X = in + S.off[sl];
    float* __restrict__ C = out + S.off[sl];
    __shared__ float Xs[kSpecTK][64 + 4];
    __shared__ float Bs[kSpecTK][64 + 4];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    float acc[4][4] = {};
    for (int i0 = 0; i0 < m; i0 += kSpecTK) {
        for (int e = tid; e < kSpecTK * 64; e += 256) {
            const int ll = e >> 4, ii = e & 15, l = l0 + ll, i = i0 + ii;  // 16 consecutive i per line
            Xs[ii][ll] = (i < m && l < L) ? X[(int64_t)l * m + i] : 0.f;
            const int kk = e & 63, ib = e >> 6, k = k0 + kk, i2 = i0 + ib;
            Bs[ib][kk] = (i2 < m && k < m) ? (INV ? Q[(int64_t)k * m + i2] : Q[(int64_t)i2 * m + k]) : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int ii = 0; ii < kSpecTK; ++ii) {
            const float4 a = *reinterpret_cast<const float4*>(&Xs[ii][ty * 4]);
            const float4 bv = *reinterpret_cast<const float4*>(&Bs[ii][tx * 4]);
            const float av[4] = {a.x, a.y, a.z, a.w}, bs[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fmaf(av[r], bs[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int l = l0 + ty * 4 + r;
        if (l >= L) continue;
        float lxy = 0.f;
        if (SCALE) {
            const int ix = l / my, iy = l - ix * my;
            lxy = S.lam[0][cx][ix] * S.scale[0] + S.lam[1][cy][iy] * S.scale[1];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int k = k0 + tx * 4 + c;
            if (k >= m) continue;
            float v = acc[r][c];
            if (SCALE) {
                const float lam = lxy + S.lam[2][cz][k] * S.scale[2];
                v = lam > 0.f ? v / lam : 0.f;
            }
            C[(int64_t)l * m + k] = v;
        }
    }
}

// z = the packed sub-lattices on the interior, 0 on the boundary layer;
// reduces r.z over the stored values, and the last block updates the CG
// (cg_after_precond).
template <typename V>
__global__ void __launch_bounds__(256) spec_scatter_kernel(const float* __restrict__ W, V* __restrict__ zo,
                                                           const V* __restrict__ r, Grid g, SubLat S,
                                                           double* part_sum, CgState* st) {
    if (st->stop) return;
    __shared__ double red[16];
    double s = 0.0;
    SW_FOR_EACH_VOXEL(g, {
        V zv = (V)0;
        if (g.interior(x, y, z)) {
            const int cx = (x - 1) & 1, cy = (y - 1) & 1, cz = (z - 1) & 1;
            const int my = S.m[1][cy], mz = S.m[2][cz];
            const int64_t j = S.off[4 * cx + 2 * cy + cz] + ((int64_t)((x - 1) >> 1) * my + ((y - 1) >> 1)) * mz +
                              ((z - 1) >> 1);
            zv = (V)W[j];
            s += (double)r[i] * (double)zv;
        }
        zo[i] = zv;
    });
    s = block_sum(s, red);
    if (publish_slot(part_sum, nullptr, blockIdx.x, s, 0.0, &st->counter, gridDim.x)) {
        double rz, unused;
        reduce_slots(part_sum, nullptr, gridDim.x, red, &rz, &unused);
        if (threadIdx.x == 0 && threadIdx.y == 0) cg_after_precond(st, rz);
    }
}

}  // namespace

// z = A^+ r on stream s (float32 vectors). W1 and W2 hold the packed interior
// (nx-2)(ny-2)(nz-2) floats each; part_sum has gridDim slots of red_blocks.
template <typename V>
inline void spectral_apply(const V* r, V* z, float* W1, float* W2, const Grid& g, const SubLat& S, unsigned red_blocks,
                           double* part_sum, CgState* st, hipStream_t s) {
    int mmax[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) mmax[a] = S.m[a][0] > S.m[a][1] ? S.m[a][0] : S.m[a][1];
    const dim3 blk(256);
    hipLaunchKernelGGL(spec_gather_kernel<V>, dim3(red_blocks), row_block(), 0, s, r, W1, g, S, st);
    // x: L = my mz, one batch
    const dim3 gx((unsigned)((mmax[1] * mmax[2] + 63) / 64), (unsigned)((mmax[0] + 63) / 64), 8);
    // y: L = mz, a batch per x point
    const dim3 gy((unsigned)((mmax[2] + 63) / 64), (unsigned)(mmax[0] * ((mmax[1] + 63) / 64)), 8);
    // z: k tiles x line tiles
    const dim3 gz((unsigned)((mmax[2] + 63) / 64), (unsigned)((mmax[0] * mmax[1] + 63) / 64), 8);
    hipLaunchKernelGGL((spec_rows_kernel<0, false>), gx, blk, 0, s, W1, W2, S, st);
    hipLaunchKernelGGL((spec_rows_kernel<1, false>), gy, blk, 0, s, W2, W1, S, st);
    hipLaunchKernelGGL((spec_cols_kernel<false, true>), gz, blk, 0, s, W1, W2, S, st);
    hipLaunchKernelGGL((spec_cols_kernel<true, false>), gz, blk, 0, s, W2, W1, S, st);
    hipLaunchKernelGGL((spec_rows_kernel<1, true>), gy, blk, 0, s, W1, W2, S, st);
    hipLaunchKernelGGL((spec_rows_kernel<0, true>), gx, blk, 0, s, W2, W1, S, st);
    hipLaunchKernelGGL(spec_scatter_kernel<V>, dim3(red_blocks), row_block(), 0, s, W1, z, r, g, S, part_sum, st);
}

}  // namespace sw
