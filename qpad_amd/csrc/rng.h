// rng.h -- counter-based Gaussian draws shared by the injection kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qpad {

// Philox4x32-10 (Salmon et al., SC'11) and the Box-Muller pair built on it;
// the same stream as the stand-in for rand_norm in the oracle.
__device__ inline void philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                  uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__device__ inline void normal4(uint64_t seed, uint64_t stream, uint32_t sub, double out[4])
{
    const double pi = 3.141592653589793;
    for (int h = 0; h < 2; h++) {
        uint32_t o[4];
        philox4x32((uint32_t)stream, (uint32_t)(stream >> 32), sub + (uint32_t)h, 0x51ED2D0Bu,
                   (uint32_t)seed, (uint32_t)(seed >> 32), o);
        double u1 = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
        double u2 = ((double)(o[2] >> 5) * 67108864.0 + (double)(o[3] >> 6)) * (1.0 / 9007199254740992.0);
        double theta = 2.0 * pi * u1;
        double amp = sqrt(-2.0 * log(1.0 - u2));
        out[2 * h] = amp * cos(theta);
        out[2 * h + 1] = amp * sin(theta);
    }
}

}  // namespace qpad
