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

// One exchange step of the reduce-scatter.  Level L (0..4) holds
// 16 NB >> L values per lane; step L -> L + 1 pairs value i with value
// i + half (half = 8 NB >> L) as described above and keeps one sum of each
// pair.  lo, hi: the pair's values of this lane.
template <int L>
__device__ __forceinline__ double reduce_pair(double lo, double hi, int lane)
{
    static_assert(L >= 0 && L < 4, "four halving steps");
    if constexpr (L == 0) {
        return swap32_sum(lo, hi);
    } else if constexpr (L == 1) {
        return swap16_sum(lo, hi);
    } else if constexpr (L == 2) {
        const bool up = (lane & 8) != 0;
        return (up ? hi : lo) + dpp<kDppRowMirror>(up ? lo : hi);
    } else {
        const bool up = (lane & 4) != 0;
        return (up ? hi : lo) + dpp<kDppRowHalfMirror>(up ? lo : hi);
    }
}

// Steps FROM .. 3 and the quad butterfly, in place on the 16 NB >> FROM
// values a lane holds at level FROM.  On return acc[0 .. NB) of lane l hold
// the 64-lane sums of values [(l >> 2) * NB, (l >> 2 + 1) * NB) of level 0.
// Requires all 64 lanes active.
template <int NB, int FROM>
__device__ __forceinline__ void wave_reduce_from(double (&acc)[(16 * NB) >> FROM])
{
    const int lane = threadIdx.x & 63;
    if constexpr (FROM < 4) {
        constexpr int half = (8 * NB) >> FROM;
#pragma unroll
        for (int i = 0; i < half; i++) acc[i] = reduce_pair<FROM>(acc[i], acc[half + i], lane);
        wave_reduce_from<NB, FROM + 1>(*reinterpret_cast<double(*)[(16 * NB) >> (FROM + 1)]>(&acc[0]));
    } else {
#pragma unroll
        for (int i = 0; i < NB; i++) {
            acc[i] = acc[i] + dpp<kDppQuadXor1>(acc[i]);
            acc[i] = acc[i] + dpp<kDppQuadXor2>(acc[i]);
        }
    }
}

// the whole reduce-scatter of 16 NB values per lane
template <int NB>
__device__ __forceinline__ void wave_reduce_scatter(double (&acc)[16 * NB])
{
    wave_reduce_from<NB, 0>(acc);
}

// Level-L value i of values generated on the fly: t0(v), v < 16 NB, is a
// lane's level-0 value v.  Only the values feeding value i are live, so a
// lane can add a pass's contributions into level-L accumulators without
// holding all 16 NB of them.
template <int NB, int L, class T0>
__device__ __forceinline__ double reduced_value(const T0 &t0, int i, int lane)
{
    if constexpr (L == 0) {
        return t0(i);
    } else {
        constexpr int half = (8 * NB) >> (L - 1);
        return reduce_pair<L - 1>(reduced_value<NB, L - 1>(t0, i, lane),
                                  reduced_value<NB, L - 1>(t0, half + i, lane), lane);
    }
}

}  // namespace qpad
