// Acoustic strain from a recorded pressure field on gfx950
// (acoustic_strain.py:75-277): epsilon = -p / K, K = rho c^2, over the gated
// samples of an (n_vox, nt) record with time fastest.
//
//  * volumetric, peak / rms / mean and gradient give S lanes per voxel over
//    the gated samples, so loads and stores are coalesced runs along t; voxel
//    coordinates and the bulk moduli are computed once per voxel.
//  * volumetric and gradient write (voxel, gated sample) outputs; the
//    gradient reads the +-1 neighbour records along x, y, z (in a tiled
//    visiting order that keeps neighbours close in time).
//  * peak / rms / mean: max |p| and float64 sums of p^2 and p, four samples
//    in flight per lane, then a butterfly over the segment.
//  * gradient statistics: an x march over (y, z) tiles that stages the
//    float64 strain of three planes in LDS, one window of samples at a time;
//    a thread per voxel keeps the maxima and the in-order sum of |grad|^2.
//
// Contraction is off and each expression keeps the reference's order, so
// the volumetric strain, the gradient, the peak and the gradient statistics
// are bit-identical to NumPy; rms and mean differ only by the rounding of
// NumPy's own float32 pairwise sums (the sums here are float64).
#include <cmath>

#include "sw_common.h"

#pragma clang fp contract(off)

namespace sw {
namespace {

template <typename T>
__device__ __forceinline__ T bulk(const T* __restrict__ rho, const T* __restrict__ c, int64_t v) {
    const T cv = c[v];
    return rho[v] * (cv * cv);  // rho_map * c_map**2 (acoustic_strain.py:38)
}

__device__ __forceinline__ int64_t gate_sample(const int32_t* __restrict__ gate, int64_t g0, int64_t k) {
    return gate ? (int64_t)gate[k] : g0 + k;
}

// The per-voxel kernels below give S lanes (a power of two, S <= 64) to each
// voxel; a wave holds 64/S voxels and walks the voxels grid-stride. Lane sl
// of a voxel's segment handles gated samples sl, sl + S, ... (a coalesced
// run along t per load).
struct Seg {
    int S, seg, sl;
    int64_t first, stride;  // voxel group of this wave, and the wave stride in groups
};

__device__ __forceinline__ Seg seg_of(int S) {
    Seg q;
    const int lane = threadIdx.x & 63;
    q.S = S;
    q.seg = lane / S;
    q.sl = lane % S;
    q.first = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    q.stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    return q;
}

// Visiting order of the two kernels that read a voxel's six neighbours: 32 x
// 32 (y, z) tiles, each walked one x plane at a time (the tile's 1024 columns
// of plane x, then plane x + 1). A voxel's neighbours are then visited within
// about one tile plane of it, except across tile edges, so their records are
// still in L2 (a record of n_gate samples is far too long for the natural
// order, where the x neighbours are a whole plane of records apart). Each XCD
// (blocks b with b % 8 equal, for a grid of a multiple of 8 blocks) walks a
// contiguous range of the order. Positions past the grid in y or z are skipped.
constexpr int kOrderTile = 32;

struct Order {
    int ntz;
    int64_t groups;                // voxel groups of 64/S positions
    int64_t first, stride, end;  // this wave's walk over the groups
};

__device__ __forceinline__ Order order_walk(const Grid& g, int S) {
    Order w;
    const int nty = (g.ny + kOrderTile - 1) / kOrderTile;
    w.ntz = (g.nz + kOrderTile - 1) / kOrderTile;
    w.groups = (int64_t)nty * w.ntz * g.nx * (kOrderTile * kOrderTile / (64 / S));
    const unsigned nb = gridDim.x, wpb = blockDim.x >> 6, b = blockIdx.x, wave = threadIdx.x >> 6;
    if ((nb & 7u) == 0u) {
        const unsigned xcd = b & 7u;
        w.first = w.groups * xcd / 8 + (int64_t)(b >> 3) * wpb + wave;
        w.stride = (int64_t)(nb >> 3) * wpb;
        w.end = w.groups * (xcd + 1) / 8;
    } else {
        w.first = (int64_t)b * wpb + wave;
        w.stride = (int64_t)nb * wpb;
        w.end = w.groups;
    }
    return w;
}

// Position o of the order: its voxel, and whether it lies in the grid.
__device__ __forceinline__ bool order_voxel(int64_t o, const Grid& g, int ntz, int& x, int& y, int& z) {
    constexpr int T2 = kOrderTile * kOrderTile;
    const int i = (int)(o % T2);
    const int64_t rest = o / T2;
    x = (int)(rest % g.nx);
    const int64_t tile = rest / g.nx;
    y = (int)(tile / ntz) * kOrderTile + i / kOrderTile;
    z = (int)(tile % ntz) * kOrderTile + i % kOrderTile;
    return y < g.ny && z < g.nz;
}

inline int64_t order_positions(const Grid& g) {
    const int64_t nty = (g.ny + kOrderTile - 1) / kOrderTile, ntz = (g.nz + kOrderTile - 1) / kOrderTile;
    return nty * ntz * g.nx * kOrderTile * kOrderTile;
}

#ifndef SW_VOL_U
#define SW_VOL_U 4
#endif
constexpr int kVolU = SW_VOL_U;  // voxel groups per wave iteration

template <typename T>
__global__ void __launch_bounds__(256) volumetric_kernel(const T* __restrict__ p, int64_t n_vox, int64_t nt,
                                                         const int32_t* __restrict__ gate, int64_t g0, int64_t n_gate,
                                                         int S, const T* __restrict__ c, const T* __restrict__ rho,
                                                         T* __restrict__ out) {
    // kVolU voxel groups per iteration, so each lane has kVolU loads in flight
    const Seg q = seg_of(S);
    const int64_t gsz = 64 / S;
    for (int64_t wv = q.first; wv * gsz < n_vox; wv += kVolU * q.stride) {  // uniform in the wave
        int64_t v[kVolU];
        bool ok[kVolU];
        T K[kVolU];
#pragma unroll
        for (int u = 0; u < kVolU; ++u) {
            v[u] = (wv + u * q.stride) * gsz + q.seg;
            ok[u] = v[u] < n_vox;
            if (!ok[u]) v[u] = 0;  // read voxel 0, store nothing
            K[u] = bulk(rho, c, v[u]);
        }
        for (int64_t k = q.sl; k < n_gate; k += S) {
            const int64_t t = gate_sample(gate, g0, k);
            T x[kVolU];
#pragma unroll
            for (int u = 0; u < kVolU; ++u) x[u] = p[v[u] * nt + t];
#pragma unroll
            for (int u = 0; u < kVolU; ++u)
                if (ok[u]) out[v[u] * n_gate + k] = -x[u] / K[u];  // acoustic_strain.py:104
        }
    }
}

// The volumetric strain over a contiguous gate read and written 16 bytes per
// lane (vec16_gate, 16-byte aligned output): chunks of Vec16<T>::n samples,
// kVolU voxel groups per iteration, the loop over chunks uniform in the wave;
// the same arithmetic per sample as volumetric_kernel.
template <typename T>
__global__ void __launch_bounds__(256) volumetric_vec_kernel(const T* __restrict__ p, int64_t n_vox, int64_t nt,
                                                             int64_t g0, int64_t n_gate, int S,
                                                             const T* __restrict__ c, const T* __restrict__ rho,
                                                             T* __restrict__ out) {
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::n;
    const int64_t nch = n_gate / N, ntv = nt / N, g0v = g0 / N;
    const V* pv = reinterpret_cast<const V*>(p);
    V* ov = reinterpret_cast<V*>(out);
    const Seg q = seg_of(S);
    const int64_t gsz = 64 / S;
    for (int64_t wv = q.first; wv * gsz < n_vox; wv += kVolU * q.stride) {  // uniform in the wave
        int64_t v[kVolU];
        bool ok[kVolU];
        T K[kVolU];
#pragma unroll
        for (int u = 0; u < kVolU; ++u) {
            v[u] = (wv + u * q.stride) * gsz + q.seg;
            ok[u] = v[u] < n_vox;
            if (!ok[u]) v[u] = 0;  // read voxel 0, store nothing
            K[u] = bulk(rho, c, v[u]);
        }
        for (int64_t c0 = 0; c0 < nch; c0 += S) {  // uniform in the wave
            const int64_t ch = c0 + q.sl;
            const bool live = ch < nch;
            const int64_t cl = live ? ch : nch - 1;
            V x[kVolU];
#pragma unroll
            for (int u = 0; u < kVolU; ++u) x[u] = pv[v[u] * ntv + g0v + cl];
#pragma unroll
            for (int u = 0; u < kVolU; ++u) {
                if (!ok[u] || !live) continue;
                V r;
#pragma unroll
                for (int j = 0; j < N; ++j) r[j] = -x[u][j] / K[u];  // acoustic_strain.py:104
                __builtin_nontemporal_store(r, ov + v[u] * nch + ch);
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ T nan_max_t(T m, T x) {
    return (x > m || x != x) ? x : m;  // NaN propagates, as np.max / np.maximum
}

template <typename T>
__global__ void __launch_bounds__(256) strain_peak_kernel(const T* __restrict__ p, int64_t n_vox, int64_t nt,
                                                          const int32_t* __restrict__ gate, int64_t g0,
                                                          int64_t n_gate, int S, const T* __restrict__ c,
                                                          const T* __restrict__ rho, T* __restrict__ peak,
                                                          T* __restrict__ rms, T* __restrict__ mean) {
    const Seg q = seg_of(S);
    for (int64_t wv = q.first;; wv += q.stride) {
        const int64_t vbase = wv * (64 / S);
        if (vbase >= n_vox) break;  // uniform in the wave
        const int64_t v = vbase + q.seg;
        const bool ok = v < n_vox;
        T mx[4] = {T(0), T(0), T(0), T(0)};
        double s2[4] = {0.0, 0.0, 0.0, 0.0}, s1[4] = {0.0, 0.0, 0.0, 0.0};
        if (ok) {
            const T* pv = p + v * nt;
            int64_t k = q.sl;
            for (; k + 3 * S < n_gate; k += 4 * S) {  // four samples in flight per lane
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const T x = pv[gate_sample(gate, g0, k + u * S)];
                    mx[u] = nan_max_t(mx[u], x < T(0) ? -x : x);
                    s2[u] += (double)x * (double)x;
                    s1[u] += (double)x;
                }
            }
            for (; k < n_gate; k += S) {
                const T x = pv[gate_sample(gate, g0, k)];
                mx[0] = nan_max_t(mx[0], x < T(0) ? -x : x);
                s2[0] += (double)x * (double)x;
                s1[0] += (double)x;
            }
        }
        T m = nan_max_t(nan_max_t(mx[0], mx[1]), nan_max_t(mx[2], mx[3]));
        double a2 = (s2[0] + s2[1]) + (s2[2] + s2[3]);
        double a1 = (s1[0] + s1[1]) + (s1[2] + s1[3]);
        for (int off = S >> 1; off > 0; off >>= 1) {
            m = nan_max_t(m, __shfl_xor(m, off, kWave));
            a2 += __shfl_xor(a2, off, kWave);
            a1 += __shfl_xor(a1, off, kWave);
        }
        if (!ok || q.sl != 0) continue;
        const T K = bulk(rho, c, v);
        const double n = (double)n_gate;
        peak[v] = m / K;                        // acoustic_strain.py:143
        rms[v] = (T)std::sqrt((T)(a2 / n)) / K;  // acoustic_strain.py:146
        mean[v] = -(T)(a1 / n) / K;              // acoustic_strain.py:149
    }
}

// peak / rms / mean over a contiguous gate read 16 bytes per lane
// (vec16_gate): lane sl of a voxel's S-lane segment (kPeakLanes) reads chunks sl, sl + S,
// ... of Vec16<T>::n samples, two chunks and kPeakVU voxel groups per
// iteration (four 16-byte loads in flight per lane), unconditional loads
// (chunk clamped, values past the gate discarded). One maximum and two float64
// partials per vector element, combined in a fixed order; the maximum is the
// same in any order, NaN included.
constexpr int kPeakVU = 2;
// Lanes per voxel of the 16-byte kernel: 4, a 64-byte run of the record per
// load instruction and 16 voxels per wave. At 128 gated float32 samples, 4
// lanes took 428 us on a 160^3 x 256 record against 466 (8), 492 (2) and
// 586 (16): more voxels per wave keep more lines in flight.
constexpr int kPeakLanes = 4;

template <typename T>
__global__ void __launch_bounds__(256) strain_peak_vec_kernel(const T* __restrict__ p, int64_t n_vox, int64_t nt,
                                                              int64_t g0, int64_t n_gate, int S,
                                                              const T* __restrict__ c, const T* __restrict__ rho,
                                                              T* __restrict__ peak, T* __restrict__ rms,
                                                              T* __restrict__ mean) {
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::n;
    const Seg q = seg_of(S);
    const int64_t gsz = 64 / S;
    const int64_t nch = n_gate / N, ntv = nt / N, g0v = g0 / N;
    const V* pv = reinterpret_cast<const V*>(p);
    for (int64_t wv = q.first; wv * gsz < n_vox; wv += kPeakVU * q.stride) {  // uniform in the wave
        int64_t v[kPeakVU], base[kPeakVU];
        bool ok[kPeakVU];
#pragma unroll
        for (int g = 0; g < kPeakVU; ++g) {
            v[g] = (wv + g * q.stride) * gsz + q.seg;
            ok[g] = v[g] < n_vox;
            base[g] = (ok[g] ? v[g] : 0) * ntv + g0v;
        }
        T mx[kPeakVU][N];
        double s2[kPeakVU][N] = {}, s1[kPeakVU][N] = {};
#pragma unroll
        for (int g = 0; g < kPeakVU; ++g)
#pragma unroll
            for (int j = 0; j < N; ++j) mx[g][j] = T(0);
        for (int64_t c0 = 0; c0 < nch; c0 += 2 * S) {  // uniform in the wave
            const int64_t ca = c0 + q.sl, cb = c0 + S + q.sl;
            const bool ia = ca < nch, ib = cb < nch;
            const int64_t la = ia ? ca : nch - 1, lb = ib ? cb : nch - 1;
            V xa[kPeakVU], xb[kPeakVU];
#pragma unroll
            for (int g = 0; g < kPeakVU; ++g) {
                xa[g] = pv[base[g] + la];
                xb[g] = pv[base[g] + lb];
            }
#pragma unroll
            for (int g = 0; g < kPeakVU; ++g)
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    if (ia) {
                        const T x = xa[g][j];
                        mx[g][j] = nan_max_t(mx[g][j], x < T(0) ? -x : x);
                        s2[g][j] += (double)x * (double)x;
                        s1[g][j] += (double)x;
                    }
                    if (ib) {
                        const T x = xb[g][j];
                        mx[g][j] = nan_max_t(mx[g][j], x < T(0) ? -x : x);
                        s2[g][j] += (double)x * (double)x;
                        s1[g][j] += (double)x;
                    }
                }
        }
#pragma unroll
        for (int g = 0; g < kPeakVU; ++g) {
            T m = mx[g][0];
            double a2 = s2[g][0], a1 = s1[g][0];
#pragma unroll
            for (int j = 1; j < N; ++j) {
                m = nan_max_t(m, mx[g][j]);
                a2 += s2[g][j];
                a1 += s1[g][j];
            }
            for (int off = S >> 1; off > 0; off >>= 1) {
                m = nan_max_t(m, __shfl_xor(m, off, kWave));
                a2 += __shfl_xor(a2, off, kWave);
                a1 += __shfl_xor(a1, off, kWave);
            }
            if (!ok[g] || q.sl != 0) continue;
            const T K = bulk(rho, c, v[g]);
            const double n = (double)n_gate;
            peak[v[g]] = m / K;                        // acoustic_strain.py:143
            rms[v[g]] = (T)std::sqrt((T)(a2 / n)) / K;  // acoustic_strain.py:146
            mean[v[g]] = -(T)(a1 / n) / K;              // acoustic_strain.py:149
        }
    }
}

#ifndef SW_GRAD_U
#define SW_GRAD_U 2
#endif
constexpr int kGradU = SW_GRAD_U;  // voxel groups per wave iteration

template <typename T>
__global__ void __launch_bounds__(256) strain_gradient_kernel(const T* __restrict__ p, Grid g, int64_t nt,
                                                              const int32_t* __restrict__ gate, int64_t g0,
                                                              int64_t n_gate, int S, const T* __restrict__ c,
                                                              const T* __restrict__ rho, T h2x, T h2y, T h2z,
                                                              T* __restrict__ gx, T* __restrict__ gy,
                                                              T* __restrict__ gz) {
    const int64_t P = g.plane(), R = g.nz;
    // kGradU voxel groups per iteration, so each lane has 6 kGradU loads in flight
    const Seg q = seg_of(S);
    const Order w = order_walk(g, S);
    for (int64_t wv = w.first; wv < w.end; wv += kGradU * w.stride) {  // uniform in the wave
        int64_t v[kGradU];
        bool ok[kGradU], in[kGradU];
        T K[kGradU][6];
        const T* nb[kGradU][6];
#pragma unroll
        for (int u = 0; u < kGradU; ++u) {
            int x = 0, y = 0, z = 0;
            ok[u] = wv + u * w.stride < w.end && order_voxel((wv + u * w.stride) * (64 / S) + q.seg, g, w.ntz, x, y, z);
            v[u] = ok[u] ? g.at(x, y, z) : 0;
            in[u] = ok[u] && g.interior(x, y, z);
            // the six neighbours (x+, x-, y+, y-, z+, z-); off the interior the voxel itself (unused)
            const int64_t off[6] = {in[u] ? P : 0, in[u] ? -P : 0, in[u] ? R : 0,
                                    in[u] ? -R : 0, in[u] ? 1 : 0, in[u] ? -1 : 0};
#pragma unroll
            for (int n = 0; n < 6; ++n) {
                K[u][n] = bulk(rho, c, v[u] + off[n]);
                nb[u][n] = p + (v[u] + off[n]) * nt;
            }
        }
        for (int64_t k = q.sl; k < n_gate; k += S) {
            const int64_t t = gate_sample(gate, g0, k);
            T pv[kGradU][6];
#pragma unroll
            for (int u = 0; u < kGradU; ++u)
#pragma unroll
                for (int n = 0; n < 6; ++n) pv[u][n] = nb[u][n][t];
#pragma unroll
            for (int u = 0; u < kGradU; ++u) {
                if (!ok[u]) continue;
                const int64_t o = v[u] * n_gate + k;
                if (!in[u]) {
                    gx[o] = gy[o] = gz[o] = T(0);
                    continue;
                }
                // strain -p / K at the neighbours (acoustic_strain.py:196), gradient_center (solver.py:726-728)
                gx[o] = (-pv[u][0] / K[u][0] - -pv[u][1] / K[u][1]) / h2x;
                gy[o] = (-pv[u][2] / K[u][2] - -pv[u][3] / K[u][3]) / h2y;
                gz[o] = (-pv[u][4] / K[u][4] - -pv[u][5] / K[u][5]) / h2z;
            }
        }
    }
}

// The gradient over a contiguous gate read and written 16 bytes per lane
// (vec16_gate, and 16-byte aligned outputs): lane sl of a voxel's S-lane
// segment takes chunks sl, sl + S, ... of Vec16<T>::n samples, reading them
// from the six neighbour records and writing the three components as one
// 16-byte store each. The loop over chunks is uniform in the wave (chunk
// index clamped, stores only for chunks in the gate); the arithmetic per
// sample is the scalar kernel's, so the values are the same bit for bit.
template <typename T>
__global__ void __launch_bounds__(256) strain_gradient_vec_kernel(const T* __restrict__ p, Grid g, int64_t nt,
                                                                  int64_t g0, int64_t n_gate, int S,
                                                                  const T* __restrict__ c, const T* __restrict__ rho,
                                                                  T h2x, T h2y, T h2z, T* __restrict__ gx,
                                                                  T* __restrict__ gy, T* __restrict__ gz) {
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::n;
    const int64_t P = g.plane(), R = g.nz;
    const int64_t nch = n_gate / N, ntv = nt / N, g0v = g0 / N;
    const V* pv = reinterpret_cast<const V*>(p);
    V* ox = reinterpret_cast<V*>(gx);
    V* oy = reinterpret_cast<V*>(gy);
    V* oz = reinterpret_cast<V*>(gz);
    const Seg q = seg_of(S);
    const Order w = order_walk(g, S);
    for (int64_t wv = w.first; wv < w.end; wv += kGradU * w.stride) {  // uniform in the wave
        int64_t v[kGradU], nb[kGradU][6];
        bool ok[kGradU], in[kGradU];
        T K[kGradU][6];
#pragma unroll
        for (int u = 0; u < kGradU; ++u) {
            int x = 0, y = 0, z = 0;
            ok[u] = wv + u * w.stride < w.end && order_voxel((wv + u * w.stride) * (64 / S) + q.seg, g, w.ntz, x, y, z);
            v[u] = ok[u] ? g.at(x, y, z) : 0;
            in[u] = ok[u] && g.interior(x, y, z);
            // the six neighbours (x+, x-, y+, y-, z+, z-); off the interior the voxel itself (unused)
            const int64_t off[6] = {in[u] ? P : 0, in[u] ? -P : 0, in[u] ? R : 0,
                                    in[u] ? -R : 0, in[u] ? 1 : 0, in[u] ? -1 : 0};
#pragma unroll
            for (int n = 0; n < 6; ++n) {
                K[u][n] = bulk(rho, c, v[u] + off[n]);
                nb[u][n] = (v[u] + off[n]) * ntv + g0v;
            }
        }
        for (int64_t c0 = 0; c0 < nch; c0 += S) {  // uniform in the wave
            const int64_t ch = c0 + q.sl;
            const bool live = ch < nch;
            const int64_t cl = live ? ch : nch - 1;
            V x[kGradU][6];
#pragma unroll
            for (int u = 0; u < kGradU; ++u)
#pragma unroll
                for (int n = 0; n < 6; ++n) x[u][n] = pv[nb[u][n] + cl];
#pragma unroll
            for (int u = 0; u < kGradU; ++u) {
                if (!ok[u] || !live) continue;
                const int64_t o = v[u] * nch + ch;
                V rx, ry, rz;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    if (!in[u]) {
                        rx[j] = ry[j] = rz[j] = T(0);
                        continue;
                    }
                    // strain -p / K at the neighbours (acoustic_strain.py:196), gradient_center (solver.py:726-728)
                    rx[j] = (-x[u][0][j] / K[u][0] - -x[u][1][j] / K[u][1]) / h2x;
                    ry[j] = (-x[u][2][j] / K[u][2] - -x[u][3][j] / K[u][3]) / h2y;
                    rz[j] = (-x[u][4][j] / K[u][4] - -x[u][5][j] / K[u][5]) / h2z;
                }
                // the outputs are not read again here: non-temporal, so the neighbour records keep the L2
                __builtin_nontemporal_store(rx, ox + o);
                __builtin_nontemporal_store(ry, oy + o);
                __builtin_nontemporal_store(rz, oz + o);
            }
        }
    }
}

__device__ __forceinline__ double nan_max(double m, double x) { return (x > m || x != x) ? x : m; }

// Gradient statistics as an x march. A block owns an 8 x 32 (y, z) tile and an
// x chunk of kGXC planes; one thread per column. The gated samples are taken
// in windows of kGW, in order; for each window the block marches x over its chunk,
// staging every plane's tile plus a one-voxel halo in LDS (a ring of three
// planes), with the strain -p / K of each staged voxel and sample computed
// once, in float64 as the reference. A staging thread takes whole voxels: K
// once per voxel, and the window's samples of a contiguous, aligned gate in
// 16-byte loads. A thread then reads its voxel's six neighbour strains from
// LDS. The maxima and the in-order sum of |grad|^2 of the thread's kGXC
// voxels stay in registers across the windows, so the sum keeps the
// reference's sample order. Each record is read about 1.6 times (halo and
// the two planes before each chunk) instead of once per neighbour.
#ifndef SW_GP_WINDOW
#define SW_GP_WINDOW 8
#endif
#ifndef SW_GP_XCHUNK
#define SW_GP_XCHUNK 8
#endif
constexpr int kGTY = 8, kGTZ = 32, kGW = SW_GP_WINDOW, kGXC = SW_GP_XCHUNK;
constexpr int kGHY = kGTY + 2, kGHZ = kGTZ + 2, kGHV = kGHY * kGHZ;
constexpr int kGThreads = kGTY * kGTZ;
constexpr int kGStage = (kGHV + kGThreads - 1) / kGThreads;  // staged voxels per thread
constexpr int kGPitch = kGHV + 4;                            // doubles per sample row of a staged plane
static_assert(kGThreads == 256, "gradient-statistics block");
static_assert(kGW % 4 == 0, "gradient-statistics window: whole 16-byte runs");

struct GpLds {
    double s[3][kGW][kGPitch];  // strain of (plane slot, sample, halo'd tile voxel)
};

// vec: the gate is the contiguous range [g0, g0 + n_gate) of a record whose
// windows start on 16-byte boundaries (p aligned, nt and g0 multiples of the
// values per 16 bytes); checked on the host.
template <typename T>
__global__ void __launch_bounds__(kGThreads) strain_gradient_peak_kernel(
    const T* __restrict__ p, Grid g, int64_t nt, const int32_t* __restrict__ gate, int64_t g0, int64_t n_gate,
    const T* __restrict__ c, const T* __restrict__ rho, double h2x, double h2y, double h2z, int nty, int ntz, int vec,
    double* __restrict__ peak_mag, double* __restrict__ peak_gx, double* __restrict__ peak_gy,
    double* __restrict__ peak_gz, double* __restrict__ rms_mag) {
    __shared__ GpLds L;
    // tile of this block: z tiles fastest, then y, then x chunk; XCD-aware, so
    // the tiles one XCD runs at once are neighbours and share their halos in L2
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);
    const int tzi = (int)(b % (unsigned)ntz);
    const unsigned rest = b / (unsigned)ntz;
    const int tyi = (int)(rest % (unsigned)nty);
    const int x0 = (int)(rest / (unsigned)nty) * kGXC, x1 = min(x0 + kGXC, g.nx);
    const int y0 = tyi * kGTY, z0 = tzi * kGTZ;
    const int tid = threadIdx.x, y = y0 + tid / kGTZ, z = z0 + tid % kGTZ;
    const bool col = y < g.ny && z < g.nz;
    const int64_t P = g.plane();
    const int hv = (tid / kGTZ + 1) * kGHZ + tid % kGTZ + 1;  // this column in the halo'd tile
    auto slot = [&](int x) { return (x - x0 + 3) % 3; };
    // The voxels this thread stages: halo'd tile voxels tid and tid + 256;
    // soff is the voxel's offset in a plane, -1 outside the grid or the tile
    // (such a voxel reads voxel 0 of the plane and is not committed).
    int32_t soff[kGStage];
#pragma unroll
    for (int i = 0; i < kGStage; ++i) {
        const int h = tid + i * kGThreads;
        const int yy = y0 - 1 + h / kGHZ, zz = z0 - 1 + h % kGHZ;
        const bool ok = h < kGHV && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz;
        soff[i] = ok ? yy * g.nz + zz : -1;  // planes hold fewer than 2^31 voxels (checked on the host)
    }
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    struct Fetch {
        T p[kGStage][kGW], r[kGStage], c[kGStage];
    };
    // Staging of plane xs (samples k0 .. k0 + kGW - 1) in two halves: fetch
    // issues every load, commit divides and writes LDS, so the loads of plane
    // x + 2 are in flight while plane x is computed. Past the gate's end
    // (the last, partial window) a voxel reads its last gated sample again.
    auto fetch = [&](int xs, int64_t k0, bool full, Fetch& f) {
        if (xs < 0 || xs >= g.nx) return;  // block-uniform
        int64_t base = (int64_t)xs * P;
        // Opaque to the optimiser (an empty asm, no instruction): otherwise it
        // hoists each unrolled plane's addresses out of the window loop and
        // runs out of registers.
        asm volatile("" : "+s"(base));
#pragma unroll
        for (int i = 0; i < kGStage; ++i) {
            const int64_t v = base + (soff[i] >= 0 ? soff[i] : 0);
            f.r[i] = rho[v];
            f.c[i] = c[v];
            const T* pv = p + v * nt;
            if (vec && full) {  // block-uniform
                const V* q = reinterpret_cast<const V*>(pv + g0 + k0);
#pragma unroll
                for (int w = 0; w < kGW / VN; ++w) {
                    const V t = q[w];
#pragma unroll
                    for (int e = 0; e < VN; ++e) f.p[i][w * VN + e] = t[e];
                }
            } else {
#pragma unroll
                for (int u = 0; u < kGW; ++u)
                    f.p[i][u] = pv[gate_sample(gate, g0, min<int64_t>(k0 + u, n_gate - 1))];
            }
        }
    };
    auto commit = [&](int xs, const Fetch& f) {
        if (xs < 0 || xs >= g.nx) return;  // block-uniform
        double(*dst)[kGPitch] = L.s[slot(xs)];
#pragma unroll
        for (int i = 0; i < kGStage; ++i) {
            if (soff[i] < 0) continue;
            const int h = tid + i * kGThreads;
            // -p / K in float64, K = rho c^2 promoted as float64 / float32 in NumPy (acoustic_strain.py:258-262)
            const double K = (double)(f.r[i] * (f.c[i] * f.c[i]));
#pragma unroll
            for (int u = 0; u < kGW; ++u) dst[u][h] = -(double)f.p[i][u] / K;
        }
    };
    double pm[kGXC], px[kGXC], py[kGXC], pz[kGXC], sum[kGXC];
#pragma unroll
    for (int j = 0; j < kGXC; ++j) pm[j] = px[j] = py[j] = pz[j] = sum[j] = 0.0;
    Fetch f;
    for (int64_t k0 = 0; k0 < n_gate; k0 += kGW) {
        const int nw = (int)min<int64_t>(kGW, n_gate - k0);
        const bool full = nw == kGW;
        fetch(x0 - 1, k0, full, f);
        commit(x0 - 1, f);
        fetch(x0, k0, full, f);
        commit(x0, f);
        fetch(x0 + 1, k0, full, f);
#pragma unroll
        for (int j = 0; j < kGXC; ++j) {
            const int x = x0 + j;
            if (x >= x1) break;  // block-uniform
            commit(x + 1, f);
            if (j + 1 < kGXC && x + 1 < x1) fetch(x + 2, k0, full, f);
            __syncthreads();
            if (col && g.interior(x, y, z)) {
                const double(*sp)[kGPitch] = L.s[slot(x + 1)];
                const double(*sc)[kGPitch] = L.s[slot(x)];
                const double(*sm)[kGPitch] = L.s[slot(x - 1)];
                for (int u = 0; u < nw; ++u) {
                    // gradient_center of the strain (solver.py:726-728), acoustic_strain.py:262-270
                    const double ax = (sp[u][hv] - sm[u][hv]) / h2x;
                    const double ay = (sc[u][hv + kGHZ] - sc[u][hv - kGHZ]) / h2y;
                    const double az = (sc[u][hv + 1] - sc[u][hv - 1]) / h2z;
                    const double m = std::sqrt(ax * ax + ay * ay + az * az);
                    pm[j] = nan_max(pm[j], m);
                    px[j] = nan_max(px[j], std::fabs(ax));
                    py[j] = nan_max(py[j], std::fabs(ay));
                    pz[j] = nan_max(pz[j], std::fabs(az));
                    sum[j] += m * m;
                }
            }
            __syncthreads();  // the next commit overwrites the slot of plane x - 1
        }
    }
    if (!col) return;
#pragma unroll
    for (int j = 0; j < kGXC; ++j) {
        const int x = x0 + j;
        if (x >= x1) break;
        const int64_t v = g.at(x, y, z);
        peak_mag[v] = pm[j];
        peak_gx[v] = px[j];
        peak_gy[v] = py[j];
        peak_gz[v] = pz[j];
        rms_mag[v] = std::sqrt(sum[j] / (double)(n_gate > 0 ? n_gate : 1));  // acoustic_strain.py:272
    }
}

int check_record(int64_t n_vox, int64_t nt, const void* gate, int64_t g0, int64_t n_gate) {
    SW_REQUIRE(n_vox >= 1 && nt >= 1, SW_ERR_ARG, "n_vox (%lld) and nt (%lld) must be positive", (long long)n_vox,
               (long long)nt);
    SW_REQUIRE(n_gate >= 0, SW_ERR_ARG, "n_gate must be >= 0");
    SW_REQUIRE(gate || (g0 >= 0 && g0 + n_gate <= nt), SW_ERR_ARG, "gate range [%lld, %lld) outside [0, %lld)",
               (long long)g0, (long long)(g0 + n_gate), (long long)nt);
    return SW_OK;
}

int gate_lanes(int64_t n_gate, int max_lanes) {
    int s = 1;
    while (s < n_gate && s < max_lanes) s <<= 1;
    return s;
}

unsigned segment_blocks(int64_t n_vox, int S) {
    const int64_t waves = (n_vox + (64 / S) - 1) / (64 / S);
    int64_t blocks = (waves + 3) / 4;
    const int64_t cap = (int64_t)device_cu_count() * 32;
    if (blocks > cap) blocks = cap;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int sw_acoustic_strain_volumetric(int dtype, const void* p, int64_t n_vox, int64_t nt, const int32_t* gate,
                                  int64_t gate_start, int64_t n_gate, const void* c, const void* rho, void* out,
                                  void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_record(n_vox, nt, gate, gate_start, n_gate))) return rc;
    SW_REQUIRE(p && c && rho && (out || n_gate == 0), SW_ERR_ARG, "null pointer");
    if (n_gate == 0) return SW_OK;
    hipStream_t s = (hipStream_t)stream;
    SW_DISPATCH(dtype, T, {
        if (!gate && (uintptr_t)out % 16 == 0 && vec16_gate<T>(p, nt, gate_start, n_gate)) {
            // 16 lanes per voxel: 821 us on a 160^3 x 256 float32 record at 128 gated samples, against
            // 834 (8), 874 (32) and 1209 (4)
            const int S = gate_lanes(n_gate / Vec16<T>::n, 16);
            hipLaunchKernelGGL(volumetric_vec_kernel<T>, dim3(segment_blocks(n_vox, S)), dim3(256), 0, s,
                               (const T*)p, n_vox, nt, gate_start, n_gate, S, (const T*)c, (const T*)rho, (T*)out);
        } else {
            const int S = gate_lanes(n_gate, 64);
            hipLaunchKernelGGL(volumetric_kernel<T>, dim3(segment_blocks(n_vox, S)), dim3(256), 0, s, (const T*)p,
                               n_vox, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho, (T*)out);
        }
    });
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_acoustic_strain_peak(int dtype, const void* p, int64_t n_vox, int64_t nt, const int32_t* gate,
                            int64_t gate_start, int64_t n_gate, const void* c, const void* rho, void* peak, void* rms,
                            void* mean, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_record(n_vox, nt, gate, gate_start, n_gate))) return rc;
    SW_REQUIRE(p && c && rho && peak && rms && mean, SW_ERR_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    SW_DISPATCH(dtype, T, {
        if (!gate && vec16_gate<T>(p, nt, gate_start, n_gate)) {
            const int S = gate_lanes(n_gate / Vec16<T>::n, kPeakLanes);
            hipLaunchKernelGGL(strain_peak_vec_kernel<T>, dim3(segment_blocks(n_vox, S)), dim3(256), 0, s,
                               (const T*)p, n_vox, nt, gate_start, n_gate, S, (const T*)c, (const T*)rho, (T*)peak,
                               (T*)rms, (T*)mean);
        } else {
            // 8 lanes on 4-byte samples: 597 us on an unaligned gate of 126 samples of a 160^3 x 256 float32
            // record, against 775 (16), 767 (4) and 1204 (32)
            const int S = gate_lanes(n_gate, 8);
            hipLaunchKernelGGL(strain_peak_kernel<T>, dim3(segment_blocks(n_vox, S)), dim3(256), 0, s,
                               (const T*)p, n_vox, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho,
                               (T*)peak, (T*)rms, (T*)mean);
        }
    });
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_acoustic_strain_gradient(int dtype, const void* p, int64_t nx, int64_t ny, int64_t nz, int64_t nt,
                                const int32_t* gate, int64_t gate_start, int64_t n_gate, const void* c,
                                const void* rho, double dx, double dy, double dz, void* gx, void* gy, void* gz,
                                void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    if ((rc = check_record(nx * ny * nz, nt, gate, gate_start, n_gate))) return rc;
    SW_REQUIRE(p && c && rho && ((gx && gy && gz) || n_gate == 0), SW_ERR_ARG, "null pointer");
    if (n_gate == 0) return SW_OK;
    hipStream_t s = (hipStream_t)stream;
    const Grid g{(int)nx, (int)ny, (int)nz};
    // 2.0 * h in float64, then the storage type: NumPy 2 casts the Python scalar to the array's dtype
    SW_DISPATCH(dtype, T, {
        const bool out16 = (uintptr_t)gx % 16 == 0 && (uintptr_t)gy % 16 == 0 && (uintptr_t)gz % 16 == 0;
        if (!gate && out16 && vec16_gate<T>(p, nt, gate_start, n_gate)) {
            // 8 lanes per voxel: 2715 us on a 160^3 x 256 float32 record at 128 gated samples, against
            // 2891 (16), 3062 (32) and 3985 (4)
            const int S = gate_lanes(n_gate / Vec16<T>::n, 8);
            hipLaunchKernelGGL(strain_gradient_vec_kernel<T>, dim3(segment_blocks(order_positions(g), S)), dim3(256),
                               0, s, (const T*)p, g, nt, gate_start, n_gate, S, (const T*)c, (const T*)rho,
                               (T)(2.0 * dx), (T)(2.0 * dy), (T)(2.0 * dz), (T*)gx, (T*)gy, (T*)gz);
        } else {
            const int S = gate_lanes(n_gate, 64);
            hipLaunchKernelGGL(strain_gradient_kernel<T>, dim3(segment_blocks(order_positions(g), S)), dim3(256), 0,
                               s, (const T*)p, g, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho,
                               (T)(2.0 * dx), (T)(2.0 * dy), (T)(2.0 * dz), (T*)gx, (T*)gy, (T*)gz);
        }
    });
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_acoustic_strain_gradient_peak(int dtype, const void* p, int64_t nx, int64_t ny, int64_t nz, int64_t nt,
                                     const int32_t* gate, int64_t gate_start, int64_t n_gate, const void* c,
                                     const void* rho, double dx, double dy, double dz, double* peak_mag,
                                     double* peak_gx, double* peak_gy, double* peak_gz, double* rms_mag,
                                     void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_grid(nx, ny, nz))) return rc;
    if ((rc = check_record(nx * ny * nz, nt, gate, gate_start, n_gate))) return rc;
    SW_REQUIRE(p && c && rho && peak_mag && peak_gx && peak_gy && peak_gz && rms_mag, SW_ERR_ARG, "null pointer");
    SW_REQUIRE(ny * nz < (int64_t(1) << 31), SW_ERR_ARG, "a plane of %lld voxels is too large", (long long)(ny * nz));
    hipStream_t s = (hipStream_t)stream;
    const Grid g{(int)nx, (int)ny, (int)nz};
    const int nty = (g.ny + kGTY - 1) / kGTY, ntz = (g.nz + kGTZ - 1) / kGTZ, ntx = (g.nx + kGXC - 1) / kGXC;
    SW_DISPATCH(dtype, T, {
        constexpr int N = Vec16<T>::n;
        const int vec = !gate && (uintptr_t)p % 16 == 0 && nt % N == 0 && gate_start % N == 0;
        hipLaunchKernelGGL(strain_gradient_peak_kernel<T>, dim3((unsigned)(nty * ntz * ntx)), dim3(kGThreads), 0, s,
                           (const T*)p, g, nt, gate, gate_start, n_gate, (const T*)c, (const T*)rho, 2.0 * dx,
                           2.0 * dy, 2.0 * dz, nty, ntz, vec, peak_mag, peak_gx, peak_gy, peak_gz, rms_mag);
    });
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
