// The strain invariants at one voxel (strain.py:16-135), shared by the
// strain kernels (strain.hip) and the two-step pass that evaluates the
// in-loop invariants of its input state (march2_step.h), so both compute
// every voxel with the same operations.
#pragma once

#include "sw_common.h"

namespace sw {

// Invariants of a Voigt strain e = (xx, yy, zz, yz+zy, xz+zx, xy+yx) in the
// reference's order of operations (strain.py:110-128), in T, no contraction;
// von Mises also in double as sqrt(2/3) * (double)dev.
template <typename T>
__device__ inline void strain_invariants_of(const T* e, T& vol, T& dev, T& ms, double& vm) {
#pragma clang fp contract(off)
    vol = ((e[0] + e[1]) + e[2]) / T(3);
    const T dxx = e[0] - vol, dyy = e[1] - vol, dzz = e[2] - vol;
    const T nrm = (dxx * dxx + dyy * dyy) + dzz * dzz;
    const T shr = (e[3] * e[3] + e[4] * e[4]) + e[5] * e[5];
    dev = sqrt(nrm + T(0.5) * shr);
    const T j2 = T(0.5) * (nrm + T(0.5) * shr);
    ms = sqrt(j2 > T(0) ? j2 : T(0));
    vm = 0.816496580927726 * (double)dev;  // sqrt(2/3)
}

// Central differences f[+1] - f[-1] of the three displacement components
// along the three axes at one voxel.
template <typename T>
struct UDiff {
    T xx, yx, zx;  // d/dx of ux, uy, uz (unscaled)
    T xy, yy, zy;  // d/dy
    T xz, yz, zz;  // d/dz
};

// The in-loop invariants of one voxel: partials are the differences times
// 1/(2h) (multiplied, with contraction), zero off the interior, and the
// invariants follow strain_invariants_of. The drop-in's stand-alone strain
// kernels divide by 2h and reproduce NumPy bit for bit instead.
template <typename T>
__device__ __forceinline__ void inloop_invariants(bool interior, const UDiff<T>& d, T ihx, T ihy, T ihz, T& vol,
                                                  T& dev, T& ms, T& vm) {
#pragma clang fp contract(on)
    T e[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (interior) {
        e[0] = d.xx * ihx;
        e[1] = d.yy * ihy;
        e[2] = d.zz * ihz;
        e[3] = d.yz * ihz + d.zy * ihy;
        e[4] = d.xz * ihz + d.zx * ihx;
        e[5] = d.xy * ihy + d.yx * ihx;
    }
    double v;
    strain_invariants_of(e, vol, dev, ms, v);
    vm = (T)v;
}

}  // namespace sw
