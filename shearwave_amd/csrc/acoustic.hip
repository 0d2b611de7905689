// Acoustic strain from a recorded pressure field on gfx950
// (acoustic_strain.py:75-277): epsilon = -p / K, K = rho c^2, over the gated
// samples of an (n_vox, nt) record with time fastest.
//
//  * every kernel gives S lanes per voxel over the gated samples, so loads
//    and stores are coalesced runs along t; voxel coordinates and the bulk
//    modulus of the voxel and its neighbours are computed once per voxel.
//  * volumetric and gradient write (voxel, gated sample) outputs; the
//    gradient reads the +-1 neighbour records along x, y, z.
//  * peak / rms / mean: max |p| and float64 sums of p^2 and p, four samples
//    in flight per lane, then a butterfly over the segment.
//  * gradient statistics: float64, maxima reduced over the segment, the sum
//    of |grad|^2 taken in sample order, as the reference's running sum.
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

template <typename T>
__global__ void __launch_bounds__(256) volumetric_kernel(const T* __restrict__ p, int64_t n_vox, int64_t nt,
                                                         const int32_t* __restrict__ gate, int64_t g0, int64_t n_gate,
                                                         int S, const T* __restrict__ c, const T* __restrict__ rho,
                                                         T* __restrict__ out) {
    const Seg q = seg_of(S);
    for (int64_t wv = q.first;; wv += q.stride) {
        const int64_t vbase = wv * (64 / S);
        if (vbase >= n_vox) break;
        const int64_t v = vbase + q.seg;
        if (v >= n_vox) continue;
        const T K = bulk(rho, c, v);
        const T* pv = p + v * nt;
        T* ov = out + v * n_gate;
        for (int64_t k = q.sl; k < n_gate; k += S) ov[k] = -pv[gate_sample(gate, g0, k)] / K;  // acoustic_strain.py:104
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

template <typename T>
__global__ void __launch_bounds__(256) strain_gradient_kernel(const T* __restrict__ p, Grid g, int64_t nt,
                                                              const int32_t* __restrict__ gate, int64_t g0,
                                                              int64_t n_gate, int S, const T* __restrict__ c,
                                                              const T* __restrict__ rho, T h2x, T h2y, T h2z,
                                                              T* __restrict__ gx, T* __restrict__ gy,
                                                              T* __restrict__ gz) {
    const int64_t n_vox = g.size();
    const int64_t P = g.plane(), R = g.nz;
    const Seg q = seg_of(S);
    for (int64_t wv = q.first;; wv += q.stride) {
        const int64_t vbase = wv * (64 / S);
        if (vbase >= n_vox) break;
        const int64_t v = vbase + q.seg;
        if (v >= n_vox) continue;
        const int x = (int)(v / P), y = (int)((v / R) % g.ny), z = (int)(v % R);
        const int64_t o = v * n_gate;
        if (!g.interior(x, y, z)) {
            for (int64_t k = q.sl; k < n_gate; k += S) gx[o + k] = gy[o + k] = gz[o + k] = T(0);
            continue;
        }
        const T kxp = bulk(rho, c, v + P), kxm = bulk(rho, c, v - P);
        const T kyp = bulk(rho, c, v + R), kym = bulk(rho, c, v - R);
        const T kzp = bulk(rho, c, v + 1), kzm = bulk(rho, c, v - 1);
        const T *pxp = p + (v + P) * nt, *pxm = p + (v - P) * nt, *pyp = p + (v + R) * nt;
        const T *pym = p + (v - R) * nt, *pzp = p + (v + 1) * nt, *pzm = p + (v - 1) * nt;
        for (int64_t k = q.sl; k < n_gate; k += S) {
            const int64_t t = gate_sample(gate, g0, k);
            // strain -p / K at the neighbours (acoustic_strain.py:196), gradient_center (solver.py:726-728)
            gx[o + k] = (-pxp[t] / kxp - -pxm[t] / kxm) / h2x;
            gy[o + k] = (-pyp[t] / kyp - -pym[t] / kym) / h2y;
            gz[o + k] = (-pzp[t] / kzp - -pzm[t] / kzm) / h2z;
        }
    }
}

__device__ __forceinline__ double nan_max(double m, double x) { return (x > m || x != x) ? x : m; }

// S = 16 lanes per voxel. The maxima reduce across the segment (order-free);
// the sum of |grad|^2 is taken in sample order through the segment's lanes
// (shuffles), as the reference's running sum.
template <typename T>
__global__ void __launch_bounds__(256) strain_gradient_peak_kernel(
    const T* __restrict__ p, Grid g, int64_t nt, const int32_t* __restrict__ gate, int64_t g0, int64_t n_gate,
    int S, const T* __restrict__ c, const T* __restrict__ rho, double h2x, double h2y, double h2z,
    double* __restrict__ peak_mag, double* __restrict__ peak_gx, double* __restrict__ peak_gy,
    double* __restrict__ peak_gz, double* __restrict__ rms_mag) {
    const int64_t n_vox = g.size();
    const int64_t P = g.plane(), R = g.nz;
    const Seg q = seg_of(S);
    const int base_lane = (threadIdx.x & 63) - q.sl;  // first lane of this segment
    for (int64_t wv = q.first;; wv += q.stride) {
        const int64_t vbase = wv * (64 / S);
        if (vbase >= n_vox) break;  // uniform in the wave
        const int64_t v = vbase + q.seg;
        const bool ok = v < n_vox;
        int x = 0, y = 0, z = 0;
        if (ok) {
            x = (int)(v / P);
            y = (int)((v / R) % g.ny);
            z = (int)(v % R);
        }
        const bool in = ok && g.interior(x, y, z);
        double kxp = 1.0, kxm = 1.0, kyp = 1.0, kym = 1.0, kzp = 1.0, kzm = 1.0;
        if (in) {  // K promoted to float64, as float64 / float32 in NumPy (acoustic_strain.py:262)
            kxp = (double)bulk(rho, c, v + P);
            kxm = (double)bulk(rho, c, v - P);
            kyp = (double)bulk(rho, c, v + R);
            kym = (double)bulk(rho, c, v - R);
            kzp = (double)bulk(rho, c, v + 1);
            kzm = (double)bulk(rho, c, v - 1);
        }
        double pm = 0.0, px = 0.0, py = 0.0, pz = 0.0, sum = 0.0;
        // rounds of S samples; every lane of the segment runs every round (shuffles)
        for (int64_t k0 = 0; k0 < n_gate; k0 += S) {
            const int64_t k = k0 + q.sl;
            double m2 = 0.0;
            if (in && k < n_gate) {
                const int64_t t = gate_sample(gate, g0, k);
                auto s = [&](int64_t u, double K) { return -(double)p[u * nt + t] / K; };
                const double ax = (s(v + P, kxp) - s(v - P, kxm)) / h2x;
                const double ay = (s(v + R, kyp) - s(v - R, kym)) / h2y;
                const double az = (s(v + 1, kzp) - s(v - 1, kzm)) / h2z;
                const double m = std::sqrt(ax * ax + ay * ay + az * az);  // acoustic_strain.py:265
                pm = nan_max(pm, m);
                px = nan_max(px, std::fabs(ax));
                py = nan_max(py, std::fabs(ay));
                pz = nan_max(pz, std::fabs(az));
                m2 = m * m;
            }
            const int cnt = (int)min<int64_t>(S, n_gate - k0);
            for (int j = 0; j < cnt; ++j) sum += __shfl(m2, base_lane + j, kWave);  // sample order
        }
        for (int off = S >> 1; off > 0; off >>= 1) {
            pm = nan_max(pm, __shfl_xor(pm, off, kWave));
            px = nan_max(px, __shfl_xor(px, off, kWave));
            py = nan_max(py, __shfl_xor(py, off, kWave));
            pz = nan_max(pz, __shfl_xor(pz, off, kWave));
        }
        if (!ok || q.sl != 0) continue;
        peak_mag[v] = pm;
        peak_gx[v] = px;
        peak_gy[v] = py;
        peak_gz[v] = pz;
        rms_mag[v] = std::sqrt(sum / (double)(n_gate > 0 ? n_gate : 1));  // acoustic_strain.py:272
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
    const int S = gate_lanes(n_gate, 64);
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(volumetric_kernel<T>, dim3(segment_blocks(n_vox, S)), dim3(256), 0, s, (const T*)p,
                                   n_vox, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho, (T*)out));
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
    const int S = gate_lanes(n_gate, 16);
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(strain_peak_kernel<T>, dim3(segment_blocks(n_vox, S)), dim3(256), 0, s,
                                   (const T*)p, n_vox, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho,
                                   (T*)peak, (T*)rms, (T*)mean));
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
    const int S = gate_lanes(n_gate, 64);
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(strain_gradient_kernel<T>, dim3(segment_blocks(g.size(), S)), dim3(256), 0, s,
                                   (const T*)p, g, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho,
                                   (T)(2.0 * dx), (T)(2.0 * dy), (T)(2.0 * dz), (T*)gx, (T*)gy, (T*)gz));
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
    hipStream_t s = (hipStream_t)stream;
    const Grid g{(int)nx, (int)ny, (int)nz};
    const int S = gate_lanes(n_gate, 16);
    SW_DISPATCH(dtype, T,
                hipLaunchKernelGGL(strain_gradient_peak_kernel<T>, dim3(segment_blocks(g.size(), S)), dim3(256), 0, s,
                                   (const T*)p, g, nt, gate, gate_start, n_gate, S, (const T*)c, (const T*)rho,
                                   2.0 * dx, 2.0 * dy, 2.0 * dz, peak_mag, peak_gx, peak_gy, peak_gz, rms_mag));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
