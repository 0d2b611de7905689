// Acoustic radiation force from a recorded pressure field on gfx950
// (compute_radiation_force / compute_radiation_force_velocity,
// radiation_force.py:27-157): a time reduction over a 4-D (voxel, time)
// array with time fastest, bound by HBM streaming.
//
//  * energy: env[t] = sum_v p[v][t]^2, the input of the reference's auto-gate
//    (radiation_force.py:66-72). Each wave walks whole voxel rows (a
//    coalesced run along t per load), lanes keeping per-sample sums; float64
//    partials are summed over the block's waves and then over blocks in a
//    fixed order: deterministic.
//  * force: S lanes (a power of two up to 16) per voxel read its gated samples
//    (a contiguous gate is passed as a range, with no index loads), four
//    independent partial sums per lane, then a fixed butterfly; waves walk
//    the voxels grid-stride. The per-voxel arithmetic follows the reference's
//    order in float64 and rounds once to the storage type.
#include "sw_common.h"

namespace sw {
namespace {

constexpr int kEnergyThreads = 256;

// A wave walks whole voxel rows (grid-stride over voxels); lane l keeps the
// sums of samples tc + l + 64 u, u < 4, of a 256-sample chunk, so one voxel's
// chunk is four coalesced 256 B loads, and kEnergyV voxels are loaded per
// iteration (16 loads in flight per lane). The loads are unconditional:
// lanes past nt read the chunk's first sample and discard it (a load under a
// per-lane condition is a branch that waits for it). The block's four waves
// are then summed in wave order into the block's partial.
constexpr int kEnergyV = 4;

// VEC (records with nt % 4 == 0 and a 16-byte aligned start): lane l takes
// samples tc + 4 l + u instead of tc + l + 64 u, four consecutive samples read
// as 16-byte runs; each sample's sum still runs over the voxels in the same
// order, so the partials are the same bit for bit.
template <typename T, bool VEC>
__global__ void __launch_bounds__(kEnergyThreads) energy_partial_kernel(const T* __restrict__ p, int64_t n_vox,
                                                                        int64_t nt, double* __restrict__ part) {
    __shared__ double red[kEnergyThreads / kWave][4 * kWave];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::n;
    // sample offset (in the chunk) of this lane's value u
    auto slot = [&](int u) { return VEC ? 4 * lane + u : lane + kWave * u; };
    for (int64_t tc = 0; tc < nt; tc += 4 * kWave) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        bool tin[4];
        int64_t off[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tin[u] = tc + slot(u) < nt;
            off[u] = tin[u] ? tc + slot(u) : tc;  // with VEC, all four in or all four out (nt % 4 == 0)
        }
        auto load = [&](int64_t vv, T (&x)[4]) {
            if (VEC) {
                const V* q = reinterpret_cast<const V*>(p + vv * nt + off[0]);
#pragma unroll
                for (int w = 0; w < 4 / N; ++w) {
                    const V t = q[w];
#pragma unroll
                    for (int e = 0; e < N; ++e) x[w * N + e] = t[e];
                }
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = p[vv * nt + off[u]];
            }
        };
        int64_t v = w0;
        for (; v + (kEnergyV - 1) * n_waves < n_vox; v += kEnergyV * n_waves) {
            T x[kEnergyV][4];
#pragma unroll
            for (int j = 0; j < kEnergyV; ++j) load(v + j * n_waves, x[j]);
#pragma unroll
            for (int j = 0; j < kEnergyV; ++j)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double d = tin[u] ? (double)x[j][u] : 0.0;
                    acc[u] += d * d;
                }
        }
        for (; v < n_vox; v += n_waves) {
            T x[4];
            load(v, x);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double d = tin[u] ? (double)x[u] : 0.0;
                acc[u] += d * d;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) red[wid][slot(u)] = acc[u];
        __syncthreads();
        for (int k = threadIdx.x; k < 4 * kWave; k += blockDim.x) {
            const int64_t t = tc + k;
            if (t < nt) {
                double s = 0.0;
                for (int w = 0; w < kEnergyThreads / kWave; ++w) s += red[w][k];  // fixed order
                part[(int64_t)blockIdx.x * nt + t] = s;
            }
        }
        __syncthreads();
    }
}

// env[t] = sum over blocks of part[b][t]: one block per sample, threads
// stride over the partials in a fixed order, then a fixed block reduction.
__global__ void __launch_bounds__(256) energy_sum_kernel(const double* __restrict__ part, int64_t nb, int64_t nt,
                                                         double* __restrict__ env) {
    __shared__ double red[16];
    for (int64_t t = blockIdx.x; t < nt; t += gridDim.x) {
        double s = 0.0;
        for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) s += part[b * nt + t];
        s = block_sum(s, red);
        if (threadIdx.x == 0) env[t] = s;
        __syncthreads();
    }
}

struct EnergyShape {
    int64_t nb;
};

EnergyShape energy_shape(int64_t n_vox, int64_t nt) {
    (void)nt;
    EnergyShape e;
    // eight blocks per CU (one resident wave of blocks), at most one wave per voxel
    const int64_t waves_per_block = kEnergyThreads / kWave;
    int64_t nb = (int64_t)device_cu_count() * 8;
    const int64_t need = (n_vox + waves_per_block - 1) / waves_per_block;
    if (nb > need) nb = need;
    e.nb = nb < 1 ? 1 : nb;
    return e;
}

constexpr double kDbPerNeper = 8.685889638;  // radiation_force.py:79, 149

// Sample k of a gate: gate[k], or g0 + k for a contiguous gate (gate == nullptr).
__device__ __forceinline__ int64_t gate_at(const int32_t* __restrict__ gate, int64_t g0, int64_t k) {
    return gate ? (int64_t)gate[k] : g0 + k;
}

template <typename T, bool VEL>
__global__ void __launch_bounds__(256) radforce_kernel(const T* __restrict__ p, const T* __restrict__ vx,
                                                       const T* __restrict__ vy, const T* __restrict__ vz,
                                                       int64_t n_vox, int64_t nt, const int32_t* __restrict__ gate,
                                                       int64_t g0, int64_t n_gate, int S, const T* __restrict__ c,
                                                       const T* __restrict__ rho, const T* __restrict__ alpha,
                                                       double f0_mhz, T* __restrict__ o0, T* __restrict__ o1,
                                                       T* __restrict__ o2) {
    const int lane = threadIdx.x & 63;
    const int seg = lane / S, sl = lane % S;
    const int per_wave = 64 / S;
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const double n = (double)n_gate;
    for (int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;; wv += n_waves) {
        const int64_t vbase = wv * per_wave;
        if (vbase >= n_vox) break;  // uniform in the wave
        const int64_t v = vbase + seg;
        const bool ok = v < n_vox;
        // four independent partial sums per lane (samples sl, sl+S, sl+2S, sl+3S, ...)
        double s0[4] = {0.0, 0.0, 0.0, 0.0}, s1[4] = {0.0, 0.0, 0.0, 0.0}, s2[4] = {0.0, 0.0, 0.0, 0.0};
        if (ok) {
            const int64_t base = v * nt;
            int64_t k = sl;
            for (; k + 3 * S < n_gate; k += 4 * S) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t i = base + gate_at(gate, g0, k + u * S);
                    const double x = (double)p[i];
                    if (VEL) {
                        s0[u] += x * (double)vx[i];
                        s1[u] += x * (double)vy[i];
                        s2[u] += x * (double)vz[i];
                    } else {
                        s0[u] += x * x;
                    }
                }
            }
            for (; k < n_gate; k += S) {
                const int64_t i = base + gate_at(gate, g0, k);
                const double x = (double)p[i];
                if (VEL) {
                    s0[0] += x * (double)vx[i];
                    s1[0] += x * (double)vy[i];
                    s2[0] += x * (double)vz[i];
                } else {
                    s0[0] += x * x;
                }
            }
        }
        double a0 = (s0[0] + s0[1]) + (s0[2] + s0[3]);
        double a1 = (s1[0] + s1[1]) + (s1[2] + s1[3]);
        double a2 = (s2[0] + s2[1]) + (s2[2] + s2[3]);
        for (int off = S >> 1; off > 0; off >>= 1) {  // within the S-lane segment
            a0 += __shfl_xor(a0, off, kWave);
            if (VEL) {
                a1 += __shfl_xor(a1, off, kWave);
                a2 += __shfl_xor(a2, off, kWave);
            }
        }
        if (!ok || sl != 0) continue;
        const double cc = (double)c[v];
        const double anp = ((double)alpha[v] * f0_mhz) / kDbPerNeper * 100.0;  // radiation_force.py:78-79
        if (VEL) {
            const double coeff = 2.0 * anp / cc;  // radiation_force.py:151-155
            o0[v] = (T)(coeff * (a0 / n));
            o1[v] = (T)(coeff * (a1 / n));
            o2[v] = (T)(coeff * (a2 / n));
        } else {
            const double isppa = (a0 / n) / ((double)rho[v] * cc);  // radiation_force.py:74-75
            o0[v] = (T)(2.0 * anp * isppa / cc);                    // radiation_force.py:82
        }
    }
}

// The force pass over a contiguous gate whose samples can be read 16 bytes
// per lane (vec16_gate): lane sl of a voxel's S-lane segment reads chunks
// sl, sl + S, ... of Vec16<T>::n samples each, kForceVU voxel groups and two
// chunks per iteration, so four 16-byte loads are in flight per lane (a
// segment reads a 64-byte run of a voxel's samples per load).
// The loads are unconditional (chunk index clamped, values past the gate
// discarded). One float64 partial per vector element, summed in a fixed
// order, then the butterfly over the segment.
constexpr int kForceVU = 2;

template <typename T>
__global__ void __launch_bounds__(256) radforce_vec_kernel(const T* __restrict__ p, int64_t n_vox, int64_t nt,
                                                           int64_t g0, int64_t n_gate, int S,
                                                           const T* __restrict__ c, const T* __restrict__ rho,
                                                           const T* __restrict__ alpha, double f0_mhz,
                                                           T* __restrict__ o0) {
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::n;
    const int lane = threadIdx.x & 63;
    const int seg = lane / S, sl = lane % S;
    const int per_wave = 64 / S;
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t nch = n_gate / N, ntv = nt / N, g0v = g0 / N;
    const V* pv = reinterpret_cast<const V*>(p);
    for (int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; wv * per_wave < n_vox;
         wv += kForceVU * n_waves) {  // uniform in the wave
        int64_t v[kForceVU], base[kForceVU];
        bool ok[kForceVU];
#pragma unroll
        for (int g = 0; g < kForceVU; ++g) {
            v[g] = (wv + g * n_waves) * per_wave + seg;
            ok[g] = v[g] < n_vox;
            base[g] = (ok[g] ? v[g] : 0) * ntv + g0v;
        }
        double s[kForceVU][N] = {};
        for (int64_t c0 = 0; c0 < nch; c0 += 2 * S) {  // uniform in the wave
            const int64_t ca = c0 + sl, cb = c0 + S + sl;
            const bool ia = ca < nch, ib = cb < nch;
            const int64_t la = ia ? ca : nch - 1, lb = ib ? cb : nch - 1;
            V xa[kForceVU], xb[kForceVU];
#pragma unroll
            for (int g = 0; g < kForceVU; ++g) {
                xa[g] = pv[base[g] + la];
                xb[g] = pv[base[g] + lb];
            }
#pragma unroll
            for (int g = 0; g < kForceVU; ++g)
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const double a = ia ? (double)xa[g][j] : 0.0, b = ib ? (double)xb[g][j] : 0.0;
                    s[g][j] += a * a;
                    s[g][j] += b * b;
                }
        }
#pragma unroll
        for (int g = 0; g < kForceVU; ++g) {
            double a0 = s[g][0];
#pragma unroll
            for (int j = 1; j < N; ++j) a0 += s[g][j];
            for (int off = S >> 1; off > 0; off >>= 1) a0 += __shfl_xor(a0, off, kWave);  // within the segment
            if (!ok[g] || sl != 0) continue;
            const int64_t vg = v[g];
            const double cc = (double)c[vg];
            const double anp = ((double)alpha[vg] * f0_mhz) / kDbPerNeper * 100.0;  // radiation_force.py:78-79
            const double isppa = (a0 / (double)n_gate) / ((double)rho[vg] * cc);  // radiation_force.py:74-75
            o0[vg] = (T)(2.0 * anp * isppa / cc);                                  // radiation_force.py:82
        }
    }
}

int check_sizes(int64_t n_vox, int64_t nt) {
    SW_REQUIRE(n_vox >= 1 && nt >= 1, SW_ERR_ARG, "n_vox (%lld) and nt (%lld) must be positive", (long long)n_vox,
               (long long)nt);
    SW_REQUIRE(nt < (int64_t(1) << 31), SW_ERR_ARG, "nt too large");
    return SW_OK;
}

// Lanes per voxel: 8 (a 32 B run of float32 samples per load and eight
// voxels per wave) once the gate has that many samples, fewer below. On an
// unaligned gate of 126 samples of a 160^3 x 256 float32 record the force
// pass took 547 us with 8 lanes, against 644 (16), 835 (4) and 958 (32).
int segment_lanes(int64_t n_gate) {
    int s = 1;
    while (s < n_gate && s < 8) s <<= 1;
    return s;
}

template <typename T, bool VEL>
void launch_force(const T* p, const T* vx, const T* vy, const T* vz, int64_t n_vox, int64_t nt, const int32_t* gate,
                  int64_t g0, int64_t n_gate, const T* c, const T* rho, const T* alpha, double f0, T* o0, T* o1,
                  T* o2, hipStream_t s) {
    if constexpr (!VEL) {
        if (!gate && vec16_gate<T>(p, nt, g0, n_gate)) {
            // 4 lanes per voxel (a 64-byte run per load, 16 voxels per wave): 401 us on a 160^3 x 256
            // float32 record at 128 gated samples, against 405 (8), 474 (16) and 551 (2)
            int S = 1;
            while (S < n_gate / Vec16<T>::n && S < 4) S <<= 1;
            const int64_t waves = (n_vox + (64 / S) - 1) / (64 / S);
            int64_t blocks = (waves + 4 * kForceVU - 1) / (4 * kForceVU);
            const int64_t cap = (int64_t)device_cu_count() * 32;
            if (blocks > cap) blocks = cap;
            hipLaunchKernelGGL(radforce_vec_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n_vox, nt, g0,
                               n_gate, S, c, rho, alpha, f0 / 1e6, o0);
            return;
        }
    }
    const int S = segment_lanes(n_gate);
    const int64_t waves = (n_vox + (64 / S) - 1) / (64 / S);
    int64_t blocks = (waves + 3) / 4;
    const int64_t cap = (int64_t)device_cu_count() * 32;  // grid-stride beyond eight waves per SIMD
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL((radforce_kernel<T, VEL>), dim3((unsigned)blocks), dim3(256), 0, s, p, vx, vy, vz, n_vox, nt,
                       gate, g0, n_gate, S, c, rho, alpha, f0 / 1e6, o0, o1, o2);
}

}  // namespace
}  // namespace sw

using namespace sw;

extern "C" {

int64_t sw_pressure_energy_workspace(int64_t n_vox, int64_t nt) {
    if (n_vox < 1 || nt < 1) return -1;
    const EnergyShape e = energy_shape(n_vox, nt);
    return e.nb * nt * (int64_t)sizeof(double);
}

int sw_pressure_energy(int dtype, const void* p, int64_t n_vox, int64_t nt, double* env, void* workspace,
                       int64_t workspace_bytes, void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_sizes(n_vox, nt))) return rc;
    SW_REQUIRE(p && env, SW_ERR_ARG, "null pointer");
    const EnergyShape e = energy_shape(n_vox, nt);
    SW_REQUIRE(workspace && workspace_bytes >= e.nb * nt * (int64_t)sizeof(double), SW_ERR_WORKSPACE,
               "workspace of %lld bytes is smaller than the %lld needed", (long long)workspace_bytes,
               (long long)(e.nb * nt * (int64_t)sizeof(double)));
    hipStream_t s = (hipStream_t)stream;
    double* part = (double*)workspace;
    const bool vec = (uintptr_t)p % 16 == 0 && nt % 4 == 0;
    SW_DISPATCH(dtype, T, {
        if (vec)
            hipLaunchKernelGGL((energy_partial_kernel<T, true>), dim3((unsigned)e.nb), dim3(kEnergyThreads), 0, s,
                               (const T*)p, n_vox, nt, part);
        else
            hipLaunchKernelGGL((energy_partial_kernel<T, false>), dim3((unsigned)e.nb), dim3(kEnergyThreads), 0, s,
                               (const T*)p, n_vox, nt, part);
    });
    SW_LAUNCH_CHECK();
    const int64_t sum_blocks = nt < 65536 ? nt : 65536;
    hipLaunchKernelGGL(energy_sum_kernel, dim3((unsigned)sum_blocks), dim3(256), 0, s, part, e.nb, nt, env);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_radiation_force(int dtype, const void* p, int64_t n_vox, int64_t nt, const int32_t* gate, int64_t gate_start,
                       int64_t n_gate, const void* c, const void* rho, const void* alpha_coeff, double f0, void* out,
                       void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_sizes(n_vox, nt))) return rc;
    SW_REQUIRE(p && c && rho && alpha_coeff && out, SW_ERR_ARG, "null pointer");
    SW_REQUIRE(n_gate >= 0, SW_ERR_ARG, "n_gate must be >= 0");
    SW_REQUIRE(gate || (gate_start >= 0 && gate_start + n_gate <= nt), SW_ERR_ARG,
               "gate range [%lld, %lld) outside [0, %lld)", (long long)gate_start, (long long)(gate_start + n_gate),
               (long long)nt);
    hipStream_t s = (hipStream_t)stream;
    SW_DISPATCH(dtype, T,
                launch_force<T, false>((const T*)p, nullptr, nullptr, nullptr, n_vox, nt, gate, gate_start, n_gate,
                                       (const T*)c, (const T*)rho, (const T*)alpha_coeff, f0, (T*)out, nullptr,
                                       nullptr, s));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int sw_radiation_force_velocity(int dtype, const void* p, const void* vx, const void* vy, const void* vz,
                                int64_t n_vox, int64_t nt, const int32_t* gate, int64_t gate_start, int64_t n_gate,
                                const void* c, const void* alpha_coeff, double f0, void* bx, void* by, void* bz,
                                void* stream) {
    int rc = check_dtype(dtype);
    if (rc) return rc;
    if ((rc = check_sizes(n_vox, nt))) return rc;
    SW_REQUIRE(p && vx && vy && vz && c && alpha_coeff && bx && by && bz, SW_ERR_ARG, "null pointer");
    SW_REQUIRE(n_gate >= 0, SW_ERR_ARG, "n_gate must be >= 0");
    SW_REQUIRE(gate || (gate_start >= 0 && gate_start + n_gate <= nt), SW_ERR_ARG,
               "gate range [%lld, %lld) outside [0, %lld)", (long long)gate_start, (long long)(gate_start + n_gate),
               (long long)nt);
    hipStream_t s = (hipStream_t)stream;
    SW_DISPATCH(dtype, T,
                launch_force<T, true>((const T*)p, (const T*)vx, (const T*)vy, (const T*)vz, n_vox, nt, gate,
                                      gate_start, n_gate, (const T*)c, nullptr, (const T*)alpha_coeff, f0, (T*)bx,
                                      (T*)by, (T*)bz, s));
    SW_LAUNCH_CHECK();
    return SW_OK;
}

}  // extern "C"
