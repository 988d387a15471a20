// tridiag.h -- workgroup-wide tridiagonal solves with precomputed Thomas
// factors, register resident.
//
// Thread t of a workgroup (T = kSolveThreads threads) owns the K consecutive
// rows r = t*K .. t*K+K-1 of every system it solves.  Thomas' forward
// recurrence
//     y_r = (d_r - lo_r y_{r-1}) inv_r
// and back substitution
//     x_r = y_r - cp_r x_{r+1}
// are affine maps of the incoming value, so each thread first composes the
// maps of its K rows, a block scan of the composed maps gives every thread
// its incoming value, and each thread then re-runs Thomas' recurrences on
// its rows from that value.  Only the scans touch LDS.  NS systems are solved
// together so that their scans share the barriers.
//
// Global memory is always accessed interleaved (thread t handles rows
// t + kk*T, coalesced); a padded LDS stage transposes between that order and
// the blocked order of the solve.  The factors are stored pre-transposed as
// [kk][t] so their loads coalesce too; rows past nr are identity rows
// (lo = 0, inv = 1, cp = 0) with a zero right-hand side.
#pragma once

#include <hip/hip_runtime.h>

#include "qpad_internal.h"

namespace qpad {

// Block scan of affine maps y -> A y + B in thread order (rev = false) or
// reversed thread order (rev = true).  On return ex[s] is, for this thread,
// the B-component of the composition of the maps of all threads before it in
// scan order (the value entering its segment when the first segment starts
// from 0).  red must hold 2 * NS * (blockDim.x / 64) doubles.
template <int NS>
__device__ inline void block_affine_scan(double (&A)[NS], double (&B)[NS], double (&ex)[NS], double *red,
                                         bool rev)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        double Ap[NS], Bp[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            Ap[s] = rev ? __shfl_down(A[s], d, 64) : __shfl_up(A[s], d, 64);
            Bp[s] = rev ? __shfl_down(B[s], d, 64) : __shfl_up(B[s], d, 64);
        }
        const bool take = rev ? (lane + d < 64) : (lane >= d);
        if (take) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                B[s] = A[s] * Bp[s] + B[s];
                A[s] = A[s] * Ap[s];
            }
        }
    }
    if (lane == (rev ? 0 : 63)) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            red[(wave * NS + s) * 2] = A[s];
            red[(wave * NS + s) * 2 + 1] = B[s];
        }
    }
    __syncthreads();
    double PB[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double pa = 1.0, pb = 0.0;
        if (!rev) {
            for (int w = 0; w < wave; w++) {
                const double wa = red[(w * NS + s) * 2], wb = red[(w * NS + s) * 2 + 1];
                pb = wa * pb + wb;
                pa = wa * pa;
            }
        } else {
            for (int w = nw - 1; w > wave; w--) {
                const double wa = red[(w * NS + s) * 2], wb = red[(w * NS + s) * 2 + 1];
                pb = wa * pb + wb;
                pa = wa * pa;
            }
        }
        PB[s] = pb;
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const double inc = A[s] * PB[s] + B[s];
        double e = rev ? __shfl_down(inc, 1, 64) : __shfl_up(inc, 1, 64);
        if (lane == (rev ? 63 : 0)) e = PB[s];
        ex[s] = e;
    }
    __syncthreads();
}

// Factors pre-transposed: element (row r = t*K + kk) at [kk * T + t].
struct FactorT {
    const double *lo, *inv, *cp;
};

// this thread's factor rows of NS systems (independent of the data, so a
// caller can request them together with its right-hand sides)
template <int K, int NS>
struct FactorRows {
    double lo[NS][K], inv[NS][K], cp[NS][K];
};

template <int K, int NS>
__device__ inline void load_factors(const FactorT (&f)[NS], FactorRows<K, NS> &r)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            r.lo[s][k] = f[s].lo[k * kSolveThreads + t];
            r.inv[s][k] = f[s].inv[k * kSolveThreads + t];
            r.cp[s][k] = f[s].cp[k * kSolveThreads + t];
        }
    }
}

// Solve NS systems in place on d[s][k] (row r = threadIdx.x * K + k), with
// the factor rows already in registers.
template <int K, int NS>
__device__ inline void seg_solve_rows(const FactorRows<K, NS> &fr, double (&d)[NS][K], double *red)
{
    double A[NS], B[NS], ex[NS];
    const auto &lo = fr.lo;
    const auto &inv = fr.inv;
    const auto &cp = fr.cp;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double a = 1.0, b = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double al = -lo[s][k] * inv[s][k];
            b = al * b + d[s][k] * inv[s][k];
            a = al * a;
        }
        A[s] = a;
        B[s] = b;
    }
    block_affine_scan<NS>(A, B, ex, red, false);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double y = ex[s];
#pragma unroll
        for (int k = 0; k < K; k++) {
            y = (d[s][k] - lo[s][k] * y) * inv[s][k];
            d[s][k] = y;
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double a = 1.0, b = 0.0;
#pragma unroll
        for (int k = K - 1; k >= 0; k--) {
            const double al = -cp[s][k];
            b = al * b + d[s][k];
            a = al * a;
        }
        A[s] = a;
        B[s] = b;
    }
    block_affine_scan<NS>(A, B, ex, red, true);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double x = ex[s];
#pragma unroll
        for (int k = K - 1; k >= 0; k--) {
            x = d[s][k] - cp[s][k] * x;
            d[s][k] = x;
        }
    }
}

// Solve NS systems in place on d[s][k] (row r = threadIdx.x * K + k).
template <int K, int NS>
__device__ inline void seg_solve(const FactorT (&f)[NS], double (&d)[NS][K], double *red)
{
    FactorRows<K, NS> fr;
    load_factors<K, NS>(f, fr);
    seg_solve_rows<K, NS>(fr, d, red);
}

// ---- padded LDS stage between interleaved and blocked row order ----
template <int K>
__device__ inline int pidx(int r)
{
    return K == 1 ? r : r + r / K;
}

template <int K>
constexpr int stage_doubles()
{
    return (K + 1) * kSolveThreads;
}

// blocked rows of this thread <- stage
template <int K>
__device__ inline void stage_to_blocked(const double *st, double (&d)[K])
{
    const int r0 = threadIdx.x * K;
#pragma unroll
    for (int k = 0; k < K; k++) d[k] = st[pidx<K>(r0 + k)];
}

template <int K>
__device__ inline void blocked_to_stage(double *st, const double (&d)[K])
{
    const int r0 = threadIdx.x * K;
#pragma unroll
    for (int k = 0; k < K; k++) st[pidx<K>(r0 + k)] = d[k];
}

}  // namespace qpad
