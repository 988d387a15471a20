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

// rel and abs residuals: ||old - new||_inf and that over ||old||_inf (huge
// when ||old|| <= epsilon).  Called by every lane of a wave; nblk <= 64.
__device__ inline void conv_residuals(const double *blk, int nblk, double &rel, double &ab)
{
    const int lane = threadIdx.x & 63;
    double o = -INFINITY, d = -INFINITY;
    if (lane < nblk) {
        o = blk[2 * lane];
        d = blk[2 * lane + 1];
    }
    for (int s = 32; s > 0; s >>= 1) {
        o = fmax(o, __shfl_xor(o, s, 64));
        d = fmax(d, __shfl_xor(d, s, 64));
    }
    const double old_norm = sqrt(o);
    ab = sqrt(d);
    rel = (old_norm > 2.220446049250313e-16) ? ab / old_norm : 1.7976931348623157e308;
}

}  // namespace qpad
