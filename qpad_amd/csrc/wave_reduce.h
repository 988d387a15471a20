// wave_reduce.h -- reduce-scatter of per-lane FP64 accumulators over a
// 64-lane wave (gfx950).
//
// A wave holds V = 16 * NB values per lane (NB >= 1) and needs the 64-lane
// sum of each.  A butterfly all-reduce costs 6 exchanges per value; here each
// exchange step halves the number of values a lane still carries:
//   lanes  0-31 | 32-63  : v_permlane32_swap  (V/2 values kept)
//   rows 0,2    | 1,3    : v_permlane16_swap  (V/4)
//   row mirror  (l <-> 15-l within 16 lanes, DPP)  (V/8)
//   half mirror (l <-> 7-l within 8 lanes, DPP)    (V/16 = NB)
// followed by a butterfly over the 4 lanes of a quad (DPP quad_perm).  The
// lane pairings of the two mirror steps differ from l ^ 8 and l ^ 4 but each
// pairs the two halves of its lane group, so every value still sums all 64
// lanes exactly once.  Lane l ends with the sums of values
// [g * NB, (g + 1) * NB), g = l >> 2 (the four lanes of a quad hold the same
// sums).  The summation order is fixed, so results are run-to-run identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qpad {

__device__ __forceinline__ void split(double v, int &lo, int &hi)
{
    const uint64_t u = __double_as_longlong(v);
    lo = (int)(uint32_t)u;
    hi = (int)(uint32_t)(u >> 32);
}

__device__ __forceinline__ double join(int lo, int hi)
{
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// a (kept by lanes 0-31) and b (kept by lanes 32-63) -> kept value + partner's
__device__ __forceinline__ double swap32_sum(double a, double b)
{
    int alo, ahi, blo, bhi;
    split(a, alo, ahi);
    split(b, blo, bhi);
    const auto rl = __builtin_amdgcn_permlane32_swap((unsigned)alo, (unsigned)blo, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap((unsigned)ahi, (unsigned)bhi, false, false);
    return join((int)rl[0], (int)rh[0]) + join((int)rl[1], (int)rh[1]);
}

// a (kept by rows 0, 2) and b (kept by rows 1, 3)
__device__ __forceinline__ double swap16_sum(double a, double b)
{
    int alo, ahi, blo, bhi;
    split(a, alo, ahi);
    split(b, blo, bhi);
    const auto rl = __builtin_amdgcn_permlane16_swap((unsigned)alo, (unsigned)blo, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap((unsigned)ahi, (unsigned)bhi, false, false);
    return join((int)rl[0], (int)rh[0]) + join((int)rl[1], (int)rh[1]);
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    int lo, hi;
    split(v, lo, hi);
    return join(__builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false), __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false));
}

constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppQuadXor1 = 0xB1;   // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadXor2 = 0x4E;   // quad_perm [2, 3, 0, 1]

// In place on acc[0 .. 16 NB): on return acc[0 .. NB) of lane l hold the
// 64-lane sums of values [(l >> 2) * NB, (l >> 2 + 1) * NB).  Requires all
// 64 lanes active.
template <int NB>
__device__ __forceinline__ void wave_reduce_scatter(double (&acc)[16 * NB])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 8 * NB; i++) acc[i] = swap32_sum(acc[i], acc[8 * NB + i]);
#pragma unroll
    for (int i = 0; i < 4 * NB; i++) acc[i] = swap16_sum(acc[i], acc[4 * NB + i]);
    {
        const bool up = (lane & 8) != 0;
#pragma unroll
        for (int i = 0; i < 2 * NB; i++) {
            const double lo = acc[i], hi = acc[2 * NB + i];
            const double recv = dpp<kDppRowMirror>(up ? lo : hi);
            acc[i] = (up ? hi : lo) + recv;
        }
    }
    {
        const bool up = (lane & 4) != 0;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const double lo = acc[i], hi = acc[NB + i];
            const double recv = dpp<kDppRowHalfMirror>(up ? lo : hi);
            acc[i] = (up ? hi : lo) + recv;
        }
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
        acc[i] = acc[i] + dpp<kDppQuadXor1>(acc[i]);
        acc[i] = acc[i] + dpp<kDppQuadXor2>(acc[i]);
    }
}

}  // namespace qpad
