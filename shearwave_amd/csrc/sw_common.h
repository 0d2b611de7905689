// Shared helpers of the shearwave gfx950 kernels: error state, launch checks,
// voxel indexing and deterministic wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "shearwave.h"

namespace sw {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// Number of compute units of the current device (cached per device).
int device_cu_count();

}  // namespace sw

#define SW_TRY_HIP(expr)                                                                             \
    do {                                                                                             \
        hipError_t sw_e_ = (expr);                                                                   \
        if (sw_e_ != hipSuccess) {                                                                   \
            sw::set_error("%s: %s (%s:%d)", #expr, hipGetErrorString(sw_e_), __FILE__, __LINE__);    \
            return SW_ERR_HIP;                                                                       \
        }                                                                                            \
    } while (0)

#define SW_LAUNCH_CHECK() SW_TRY_HIP(hipGetLastError())

#define SW_REQUIRE(cond, code, ...)  \
    do {                             \
        if (!(cond)) {               \
            sw::set_error(__VA_ARGS__); \
            return (code);           \
        }                            \
    } while (0)

namespace sw {

constexpr int kWave = 64;

// Grid of (nx, ny, nz) voxels, C order, z fastest.
struct Grid {
    int nx, ny, nz;
    __host__ __device__ int64_t plane() const { return (int64_t)ny * nz; }
    __host__ __device__ int64_t size() const { return (int64_t)nx * ny * nz; }
    __host__ __device__ int64_t at(int x, int y, int z) const { return ((int64_t)x * ny + y) * nz + z; }
    __host__ __device__ bool interior(int x, int y, int z) const {
        return x > 0 && y > 0 && z > 0 && x < nx - 1 && y < ny - 1 && z < nz - 1;
    }
};

inline int check_grid(int64_t nx, int64_t ny, int64_t nz) {
    SW_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, SW_ERR_ARG, "grid (%lld,%lld,%lld) must be positive", (long long)nx,
               (long long)ny, (long long)nz);
    SW_REQUIRE(nx < (1 << 20) && ny < (1 << 20) && nz < (1 << 20), SW_ERR_ARG, "grid axis too large");
    return SW_OK;
}

// The x-marching kernels address a plane with 32-bit lane byte offsets.
constexpr int64_t kMaxMarchPlane = int64_t(1) << 28;

// A pointer that is the same in every lane, marked so (readfirstlane): the
// compiler keeps it in scalar registers and cannot fold lane offsets into it.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// T in the global address space. Every buffer the kernels address through
// at_byte is device (global) memory.
template <typename T>
using Global = __attribute__((address_space(1))) T;

// Element at a 32-bit byte offset from a base that is uniform across the wave
// (callers guarantee it): the access takes the scalar-base +
// 32-bit-vector-offset form, so no 64-bit address is formed or kept per lane.
// The rebuilt base is a global pointer: the loads and stores are global_*
// instructions, not flat ones. A flat access also counts against the LDS
// counter, so the lgkmcnt(0) wait ahead of each barrier would drain every
// prefetched load with it.
template <typename T>
__device__ __forceinline__ Global<T>& at_byte(T* base, uint32_t off) {
    using Byte = Global<std::conditional_t<std::is_const<T>::value, const char, char>>;
    const uint64_t v = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    Byte* b = reinterpret_cast<Byte*>(((uint64_t)hi << 32) | lo);
    return *reinterpret_cast<Global<T>*>(b + off);
}

// Row-structured launch used by the pointwise and small-stencil kernels:
// block (64, 4); blockIdx.x walks groups of 4 (x, y) rows; lanes walk z.
constexpr int kRowsPerBlock = 4;

inline dim3 row_block() { return dim3(kWave, kRowsPerBlock, 1); }

inline unsigned row_grid(const Grid& g, int64_t max_blocks) {
    int64_t rows = (int64_t)g.nx * g.ny;
    int64_t groups = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
    if (max_blocks > 0 && groups > max_blocks) groups = max_blocks;
    return (unsigned)(groups < 1 ? 1 : groups);
}

// Deterministic wave64 sum (fixed butterfly order).
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

// Block sum over all threads of a block of up to 1024 threads; every thread gets the result.
// `red` must hold 16 doubles of shared memory. Deterministic for a fixed block shape.
__device__ inline double block_sum(double v, double* red) {
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int nwaves = (nthreads + kWave - 1) / kWave;
    v = wave_sum(v);
    __syncthreads();
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nwaves; ++w) s += red[w];
    return s;
}

__device__ inline double block_max(double v, double* red) {
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int nwaves = (nthreads + kWave - 1) / kWave;
    v = wave_max(v);
    __syncthreads();
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < nwaves; ++w) s = fmax(s, red[w]);
    return s;
}

// Single-use streams are read with the non-temporal hint, so they do not
// evict lines that neighbouring tiles re-read from the L2 (SW_NO_NT: plain).
template <typename T>
__device__ __forceinline__ T load_once(const T* p) {
#ifndef SW_NO_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

template <typename T>
__device__ __forceinline__ T load_once(const Global<T>* p) {
#ifndef SW_NO_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

// Stores of data that is next read one launch later, after the L2 has turned
// over, carry the non-temporal hint too (SW_NO_NT: plain stores).
template <typename T>
__device__ __forceinline__ void store_out(T* p, T v) {
#ifndef SW_NO_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

template <typename T>
__device__ __forceinline__ void store_out(Global<T>* p, T v) {
#ifndef SW_NO_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Visit every voxel of g once: rows (x, y) by block-row, z by lane.
#define SW_FOR_EACH_VOXEL(g, ...)                                                       \
    do {                                                                                \
        const int64_t rows_ = (int64_t)(g).nx * (g).ny;                                 \
        for (int64_t r0_ = (int64_t)blockIdx.x * kRowsPerBlock; r0_ < rows_;            \
             r0_ += (int64_t)gridDim.x * kRowsPerBlock) {                               \
            const int64_t row_ = r0_ + threadIdx.y;                                     \
            if (row_ >= rows_) continue;                                                \
            const int x = (int)(row_ / (g).ny);                                         \
            const int y = (int)(row_ - (int64_t)x * (g).ny);                            \
            const int64_t base_ = row_ * (g).nz;                                        \
            for (int z = threadIdx.x; z < (g).nz; z += kWave) {                         \
                const int64_t i = base_ + z;                                            \
                __VA_ARGS__                                                             \
            }                                                                           \
        }                                                                               \
    } while (0)

inline unsigned flat_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    const int64_t cap = (int64_t)device_cu_count() * 16;
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

// The four strain invariants of u every few steps of sw_fdtd_run into out
// (SoA: vol, dev, max_shear, von_mises, n each, in dtype), x-marching; one
// rounding apart from sw_strain_invariants_from_u (strain.hip).
int strain_invariants_in_loop(int dtype, const void* ux, const void* uy, const void* uz, void* out, int64_t nx,
                              int64_t ny, int64_t nz, double dx, double dy, double dz, hipStream_t s);

inline int check_dtype(int dtype) {
    SW_REQUIRE(dtype == SW_F32 || dtype == SW_F64, SW_ERR_ARG, "dtype %d is not SW_F32 or SW_F64", dtype);
    return SW_OK;
}

template <typename T>
struct DType;
template <>
struct DType<float> {
    static constexpr int code = SW_F32;
};
template <>
struct DType<double> {
    static constexpr int code = SW_F64;
};

// 16 bytes of consecutive samples, loaded by one lane in one instruction:
// four float32 or two float64 values.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int n = 2;
};

// Whether a record of nt samples per voxel at p can be read in 16-byte runs
// along t over the contiguous gate [g0, g0 + n_gate).
template <typename T>
inline bool vec16_gate(const void* p, int64_t nt, int64_t g0, int64_t n_gate) {
    constexpr int N = Vec16<T>::n;
    return ((uintptr_t)p % 16 == 0) && nt % N == 0 && g0 % N == 0 && n_gate % N == 0 && n_gate > 0;
}

}  // namespace sw

#define SW_DISPATCH(dtype, T, ...)       \
    do {                                 \
        if ((dtype) == SW_F32) {         \
            using T = float;             \
            __VA_ARGS__;                 \
        } else {                         \
            using T = double;            \
            __VA_ARGS__;                 \
        }                                \
    } while (0)

