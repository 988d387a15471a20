// conv.h -- convergence_tester 'compare' (simulation_class.f03:560-599) from
// the per-workgroup maxima of the B_perp field kernel.
//
// Iteration l's field kernel writes, per workgroup b, the maxima over its
// nodes of old^2 and (old - new)^2 of the B_phi sums into blk[2b], blk[2b+1]
// (blk = the half of the buffer for l's parity).  The next iteration's
// kernels reduce them here; no cross-workgroup atomics or device-wide fences
// are needed inside the field kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace qpad {

// most workgroup maxima a wave reduces (three per lane)
constexpr int kConvMaxBlocks = 192;

// this lane's maxima over the workgroup maxima lane, lane + 64, lane + 128
// (-inf past nblk); nblk <= kConvMaxBlocks.  All loads are issued first.
__device__ inline void conv_load(const double *blk, int nblk, double &o, double &d)
{
    const int lane = threadIdx.x & 63;
    double ov[3], dv[3];
#pragma unroll
    for (int u = 0; u < 3; u++) {
        const int q = min(lane + 64 * u, nblk - 1);
        ov[u] = blk[2 * q];
        dv[u] = blk[2 * q + 1];
    }
    o = -INFINITY;
    d = -INFINITY;
#pragma unroll
    for (int u = 0; u < 3; u++) {
        if (lane + 64 * u < nblk) {
            o = fmax(o, ov[u]);
            d = fmax(d, dv[u]);
        }
    }
}

// rel and abs residuals: ||old - new||_inf and that over ||old||_inf (huge
// when ||old|| <= epsilon), from every lane's conv_load pair (whole wave)
__device__ inline void conv_finish(double o, double d, double &rel, double &ab)
{
    for (int s = 32; s > 0; s >>= 1) {
        o = fmax(o, __shfl_xor(o, s, 64));
        d = fmax(d, __shfl_xor(d, s, 64));
    }
    const double old_norm = sqrt(o);
    ab = sqrt(d);
    rel = (old_norm > 2.220446049250313e-16) ? ab / old_norm : 1.7976931348623157e308;
}

__device__ inline void conv_residuals(const double *blk, int nblk, double &rel, double &ab)
{
    double o, d;
    conv_load(blk, nblk, o, d);
    conv_finish(o, d, rel, ab);
}

}  // namespace qpad
