// field_src.h -- per-node source and post-processing expressions of QPAD's
// radial field equations, shared by the per-call field kernels and the fused
// slice kernels.  Every expression reads its inputs through an accessor
// A(c, k, i) (component c, mode array k, node i), so the same code reads a
// field array in HBM (FieldArr) or a tile staged in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include "qpad_internal.h"
#include "tridiag.h"

namespace qpad {

#define FX(p, c, k, i) (p)[((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)]

struct FieldCommon {
    int nr, ld, nm, M;
    int kt;               // rows per transposed factor array (K * kSolveThreads)
    double dr, idr, idrh, dxi, relax_fac;
    const double *fac;
};

inline FieldCommon field_common(const qpad_ctx *c)
{
    FieldCommon g;
    g.nr = c->g.nr;
    g.ld = c->g.ld;
    g.nm = c->g.nm;
    g.M = c->g.M;
    g.dr = c->g.dr;
    g.idr = 1.0 / c->g.dr;
    g.idrh = 0.5 * g.idr;
    g.dxi = c->g.dxi;
    g.relax_fac = c->g.relax_fac;
    g.fac = c->fac;
    g.kt = rows_per_thread(c->g.nr) * kSolveThreads;
    return g;
}

__device__ inline int mode_of(int k) { return (k + 1) >> 1; }
__device__ inline bool is_im(int k) { return k > 0 && (k & 1) == 0; }
__device__ inline int partner(int k) { return k == 0 ? 0 : (is_im(k) ? k - 1 : k + 1); }

// a slice field array [c][k][i] in device memory
struct FieldArr {
    const double *p;
    int nm, ld;
    __device__ __forceinline__ double operator()(int c, int k, int i) const { return p[((size_t)c * nm + k) * ld + i]; }
};

// Bz source: field_b_class.f03:241-298 (own = k, part = partner(k))
template <class A>
__device__ __forceinline__ double bz_source(const FieldCommon &g, const A &cu, int k, int i)
{
    const int nr = g.nr;
    const int m = mode_of(k);
    const double idr = g.idr, idrh = g.idrh;
    if (m == 0) {
        if (i == 1) return -2.0 * idr * cu(1, 0, 2);
        if (i == nr) {
            double ir = idr / (double)(nr - 1);
            return -idrh * (3.0 * cu(1, 0, nr) - 4.0 * cu(1, 0, nr - 1) + cu(1, 0, nr - 2)) - ir * cu(1, 0, nr);
        }
        double ir = idr / (double)(i - 1);
        return -idrh * (cu(1, 0, i + 1) - cu(1, 0, i - 1)) - ir * cu(1, 0, i);
    }
    const int p = partner(k);
    const bool im = is_im(k);
    const double dm = (double)m;
    double t, ir;
    if (i == nr) {
        ir = idr / (double)(nr - 1);
        t = -idrh * (3.0 * cu(1, k, nr) - 4.0 * cu(1, k, nr - 1) + cu(1, k, nr - 2)) - ir * cu(1, k, nr);
        return im ? t + dm * ir * cu(0, p, nr) : t - dm * ir * cu(0, p, nr);
    }
    if (i == 1) {
        if (m % 2 == 0) {
            t = -2.0 * idr * cu(1, k, 2);
            return im ? t + dm * idr * cu(0, p, 2) : t - dm * idr * cu(0, p, 2);
        }
        return 0.0;
    }
    if (i == 2 && m == 1) {
        ir = idr;
        t = -idr * (cu(1, k, 3) - cu(1, k, 2)) - ir * cu(1, k, 2);
        return im ? t + dm * ir * cu(0, p, 2) : t - dm * ir * cu(0, p, 2);
    }
    ir = idr / (double)(i - 1);
    t = -idrh * (cu(1, k, i + 1) - cu(1, k, i - 1)) - ir * cu(1, k, i);
    return im ? t + dm * ir * cu(0, p, i) : t - dm * ir * cu(0, p, i);
}

// Ez source of m = 0 at node i >= 2 (field_e_class.f03:189-209); the axis
// row (i = 1) is the discrete Gauss-law sum formed by the solve
template <class A>
__device__ __forceinline__ double ez_source_0(const FieldCommon &g, const A &cu, int i)
{
    const int nr = g.nr;
    const double idr = g.idr, idrh = g.idrh;
    if (i >= 2 && i <= nr - 2) {
        const double ir = idr / (double)(i - 1);
        return idrh * (cu(0, 0, i + 1) - cu(0, 0, i - 1)) + ir * cu(0, 0, i);
    }
    if (i == nr - 1) {
        const double ir = idr / (double)(nr - 2);
        return idrh * (cu(0, 0, nr) - cu(0, 0, nr - 2)) + ir * cu(0, 0, nr - 1);
    }
    if (i == nr) {
        const double ir = idr / (double)(nr - 1);
        return idr * (cu(0, 0, nr) - cu(0, 0, nr - 1)) + ir * cu(0, 0, nr);
    }
    return 0.0;
}

// correction of the Gauss-law axis sum (field_e_class.f03:205-209): the
// axis value is -8 (sum_{i=2}^{nr-2} src_i (i-1) - this)
template <class A>
__device__ __forceinline__ double ez_axis_edge(const FieldCommon &g, const A &cu)
{
    const int nr = g.nr;
    return g.idrh * (cu(0, 0, nr - 2) + cu(0, 0, nr - 1)) * ((double)nr - 2.5);
}

// Ez source for m > 0 (field_e_class.f03:212-268)
template <class A>
__device__ __forceinline__ double ez_source_m(const FieldCommon &g, const A &cu, int k, int i)
{
    const int nr = g.nr;
    const int m = mode_of(k), p = partner(k);
    const bool im = is_im(k);
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
    double t, ir;
    if (i == nr) {
        ir = idr / (double)(nr - 1);
        t = idrh * (3.0 * cu(0, k, nr) - 4.0 * cu(0, k, nr - 1) + cu(0, k, nr - 2)) + ir * cu(0, k, nr);
        return im ? t + dm * ir * cu(1, p, nr) : t - dm * ir * cu(1, p, nr);
    }
    if (i == 1) {
        if (m % 2 == 0) {
            t = 2.0 * idr * cu(0, k, 2);
            return im ? t + dm * idr * cu(1, p, 2) : t - dm * idr * cu(1, p, 2);
        }
        return 0.0;
    }
    if (i == 2 && m == 1) {
        ir = idr;
        t = idr * (cu(0, k, 3) - cu(0, k, 2)) + ir * cu(0, k, 2);
        return im ? t + dm * ir * cu(1, p, 2) : t - dm * ir * cu(1, p, 2);
    }
    ir = idr / (double)(i - 1);
    t = idrh * (cu(0, k, i + 1) - cu(0, k, i - 1)) + ir * cu(0, k, i);
    return im ? t + dm * ir * cu(1, p, i) : t - dm * ir * cu(1, p, i);
}

// E_perp at node i (1..nr) of mode array k from the total B at that node
// (br = B_r, bphi = B_phi) and psi (field_e_class.f03:443-504).  psik reads
// psi of mode array k; psip_i and psip_2 are psi of the partner array (the
// other of Re/Im of the same m) at nodes i and 2.  out = {E_r, E_phi}.
template <class A>
__device__ __forceinline__ void et_core(const FieldCommon &g, double br, double bphi, const A &psik,
                                        double psip_i, double psip_2, int k, int i, double out[2])
{
    const int nr = g.nr;
    const int m = mode_of(k);
    const double idr = g.idr, idrh = g.idrh;
    if (m == 0) {
        if (i == 1) { out[0] = 0.0; out[1] = 0.0; return; }
        if (i == nr)
            out[0] = bphi + idrh * (4.0 * psik(0, 0, nr - 1) - psik(0, 0, nr - 2) - 3.0 * psik(0, 0, nr));
        else
            out[0] = bphi - idrh * (psik(0, 0, i + 1) - psik(0, 0, i - 1));
        out[1] = -br;
        return;
    }
    const bool im = is_im(k);
    const double dm = (double)m;
    if (i == 1) {
        if (m == 1) {
            out[0] = bphi - idr * psik(0, k, 2);
            out[1] = im ? -br - idr * psip_2 : -br + idr * psip_2;
        } else {
            out[0] = 0.0;
            out[1] = 0.0;
        }
        return;
    }
    double ir = idr / (double)(i - 1);
    if (i == nr)
        out[0] = bphi + idrh * (4.0 * psik(0, k, nr - 1) - psik(0, k, nr - 2) - 3.0 * psik(0, k, nr));
    else
        out[0] = bphi - idrh * (psik(0, k, i + 1) - psik(0, k, i - 1));
    out[1] = im ? -br - ir * dm * psip_i : -br + ir * dm * psip_i;
}

// et_core with psi of mode array k given as the window psi(i-2 .. i+1)
// (clamped at the edges) so that a caller can load it branch-free
__device__ __forceinline__ void et_window(const FieldCommon &g, double br, double bphi, double psim2,
                                          double psim1, double psi0, double psip1, double psip_i,
                                          double psip_2, int k, int i, double out[2])
{
    const int nr = g.nr;
    const int m = mode_of(k);
    const double idr = g.idr, idrh = g.idrh;
    if (m == 0) {
        if (i == 1) { out[0] = 0.0; out[1] = 0.0; return; }
        if (i == nr) out[0] = bphi + idrh * (4.0 * psim1 - psim2 - 3.0 * psi0);
        else out[0] = bphi - idrh * (psip1 - psim1);
        out[1] = -br;
        return;
    }
    const bool im = is_im(k);
    const double dm = (double)m;
    if (i == 1) {
        if (m == 1) {
            out[0] = bphi - idr * psip1;
            out[1] = im ? -br - idr * psip_2 : -br + idr * psip_2;
        } else {
            out[0] = 0.0;
            out[1] = 0.0;
        }
        return;
    }
    double ir = idr / (double)(i - 1);
    if (i == nr) out[0] = bphi + idrh * (4.0 * psim1 - psim2 - 3.0 * psi0);
    else out[0] = bphi - idrh * (psip1 - psim1);
    out[1] = im ? -br - ir * dm * psip_i : -br + ir * dm * psip_i;
}

template <class A>
__device__ __forceinline__ void et_node(const FieldCommon &g, double br, double bphi, const A &psi, int k,
                                        int i, double out[2])
{
    const int p = partner(k);
    et_core(g, br, bphi, psi, psi(0, p, i), psi(0, p, 2), k, i, out);
}

// dJ_perp/dxi at node i of mode array k (field_src_class.f03:315-400):
// out[c] = dcu(c, k, i), c = 0 (r), 1 (phi)
template <class A1, class A2>
__device__ __forceinline__ void djdxi_node(const FieldCommon &g, const A1 &acu, const A2 &amu, int k, int i,
                                           double out[2])
{
    const int nr = g.nr;
    const int m = mode_of(k), p = partner(k);
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
    const double sg = is_im(k) ? -1.0 : 1.0;
    for (int c = 0; c < 2; c++) {
        double v;
        if (m == 0) {
            if (i == 1) v = 0.0;
            else if (i == nr) {
                double ir = idr / (double)(nr - 1);
                v = acu(c, 0, nr) + idrh * (4.0 * amu(c, 0, nr - 1) - amu(c, 0, nr - 2) - 3.0 * amu(c, 0, nr))
                    - ir * amu(c, 0, nr);
            } else {
                double ir = idr / (double)(i - 1);
                v = acu(c, 0, i) - idrh * (amu(c, 0, i + 1) - amu(c, 0, i - 1)) - ir * amu(c, 0, i);
            }
        } else {
            if (i == nr) {
                double ir = idr / (double)(nr - 1);
                double t = acu(c, k, nr) + idrh * (4.0 * amu(c, k, nr - 1) - amu(c, k, nr - 2)
                                                   - 3.0 * amu(c, k, nr)) - ir * amu(c, k, nr);
                v = sg > 0 ? t + dm * ir * amu(c + 1, p, nr) : t - dm * ir * amu(c + 1, p, nr);
            } else if (i == 1) {
                if (m == 1) {
                    double t = acu(c, k, 1) - 2.0 * idr * amu(c, k, 2);
                    v = sg > 0 ? t + dm * idr * amu(c + 1, p, 2) : t - dm * idr * amu(c + 1, p, 2);
                } else {
                    v = 0.0;
                }
            } else if (i == 2 && m == 2) {
                double ir = idr;
                double t = acu(c, k, 2) - idr * (amu(c, k, 3) - amu(c, k, 2)) - ir * amu(c, k, 2);
                v = sg > 0 ? t + dm * ir * amu(c + 1, p, 2) : t - dm * ir * amu(c + 1, p, 2);
            } else {
                double ir = idr / (double)(i - 1);
                double t = acu(c, k, i) - idrh * (amu(c, k, i + 1) - amu(c, k, i - 1)) - ir * amu(c, k, i);
                v = sg > 0 ? t + dm * ir * amu(c + 1, p, i) : t - dm * ir * amu(c + 1, p, i);
            }
        }
        out[c] = v;
    }
}

// B+ / B- sources of mode m at node i (field_b_class.f03:388-497) from
// dcu(i) (dr_ = Re, di_ = Im), cu (f2) and the current b_spe (f3):
// out = {Re B+, Im B+, Re B-, Im B-}
template <class A2, class A3>
__device__ __forceinline__ void bpm_source(const FieldCommon &g, const A2 &f2, const A3 &f3, int m, int i,
                                           const double dr_[2], const double di_[2], double out[4])
{
    const int nr = g.nr;
    const int kr = m == 0 ? 0 : 2 * m - 1, ki = 2 * m;
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
    const double rx = g.relax_fac * (idr * idr);
    if (m == 0) {
        double s1, s2;
        if (i == 1) { s1 = 0.0; s2 = 0.0; }
        else if (i == nr) {
            s1 = -dr_[1] - f3(0, 0, nr) * rx;
            s2 = dr_[0] + idrh * (3.0 * f2(2, 0, nr) - 4.0 * f2(2, 0, nr - 1) + f2(2, 0, nr - 2)) - f3(1, 0, nr) * rx;
        } else if (i == 2) {
            s1 = -dr_[1] - f3(0, 0, i) * rx;
            s2 = dr_[0] + idr * (f2(2, 0, 3) - f2(2, 0, 2)) - f3(1, 0, 2) * rx;
        } else {
            s1 = -dr_[1] - f3(0, 0, i) * rx;
            s2 = dr_[0] + idrh * (f2(2, 0, i + 1) - f2(2, 0, i - 1)) - f3(1, 0, i) * rx;
        }
        out[0] = s1;
        out[1] = 0.0;
        out[2] = s2;
        out[3] = 0.0;
        return;
    }
    double s1r, s1i, s2r, s2i;
    if (i == 1) {
        if (m == 1) {
            s1r = -dr_[1] + idr * dm * f2(2, ki, 2);
            s1i = -di_[1] - idr * dm * f2(2, kr, 2);
            s2r = dr_[0] + idr * f2(2, kr, 2);
            s2i = di_[0] + idr * f2(2, ki, 2);
        } else if (m % 2 == 0) {
            s1r = 0.0; s1i = 0.0; s2r = 0.0; s2i = 0.0;
        } else {
            s1r = idr * dm * f2(2, ki, 2);
            s1i = idr * dm * f2(2, kr, 2);
            s2r = idr * f2(2, kr, 2);
            s2i = idr * f2(2, ki, 2);
        }
    } else if (i == nr) {
        const double ir = idr / (double)(nr - 1);
        s1r = -dr_[1] + dm * f2(2, ki, nr) * ir;
        s1i = -di_[1] - dm * f2(2, kr, nr) * ir;
        s2r = dr_[0] + idrh * (3.0 * f2(2, kr, nr) - 4.0 * f2(2, kr, nr - 1) + f2(2, kr, nr - 2));
        s2i = di_[0] + idrh * (3.0 * f2(2, ki, nr) - 4.0 * f2(2, ki, nr - 1) + f2(2, ki, nr - 2));
    } else {
        const double ir = idr / (double)(i - 1);
        s1r = -dr_[1] + dm * f2(2, ki, i) * ir;
        s1i = -di_[1] - dm * f2(2, kr, i) * ir;
        s2r = dr_[0] + idrh * (f2(2, kr, i + 1) - f2(2, kr, i - 1));
        s2i = di_[0] + idrh * (f2(2, ki, i + 1) - f2(2, ki, i - 1));
    }
    const double f3r0 = f3(0, kr, i), f3r1 = f3(1, kr, i);
    const double f3i0 = f3(0, ki, i), f3i1 = f3(1, ki, i);
    out[0] = s1r - s2i - (f3r0 - f3i1) * rx;
    out[1] = s1i + s2r - (f3i0 + f3r1) * rx;
    out[2] = s1r + s2i - (f3r0 + f3i1) * rx;
    out[3] = s1i - s2r - (f3i0 - f3r1) * rx;
}

// recombination of the B+/B- solutions of mode m at node i into b_spe
// (field_b_class.f03:715-752): out = {B_r(Re), B_r(Im), B_phi(Re), B_phi(Im)}
// (for m = 0 only the Re entries are used)
__device__ __forceinline__ void bpm_recombine(int m, int i, const double s[4], double out[4])
{
    if (m == 0) {
        out[0] = (i == 1) ? 0.0 : s[0];
        out[1] = 0.0;
        out[2] = (i == 1) ? 0.0 : s[2];
        out[3] = 0.0;
        return;
    }
    double v0r = 0.5 * (s[0] + s[2]);
    double v0i = 0.5 * (s[1] + s[3]);
    double v1r = 0.5 * (s[1] - s[3]);
    double v1i = 0.5 * (-s[0] + s[2]);
    if (i == 1 && m != 1) { v0r = 0.0; v0i = 0.0; v1r = 0.0; v1i = 0.0; }
    out[0] = v0r;
    out[1] = v0i;
    out[2] = v1r;
    out[3] = v1i;
}

// ---------------------------------------------------------------------------
// Block reductions (fixed order: deterministic)
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *red)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < nw; w++) s += red[w];
    __syncthreads();
    return s;
}

__device__ inline double block_max(double v, double *red)
{
    for (int d = 32; d > 0; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < nw; w++) s = fmax(s, red[w]);
    __syncthreads();
    return s;
}

// convergence_tester 'compare' (simulation_class.f03:560-599), run by the
// last of the M + 1 B_perp workgroups of a launch to finish (all threads of
// that workgroup call it; the others return).  conv_re/conv_im hold the
// recorded sum_m |Re B_phi,m| and sum_m |Im B_phi,m| and are replaced by the
// new sums, which is the next iteration's 'record'.
template <int K>
__device__ __forceinline__ void conv_compare_last(const FieldCommon &g, const double *b_spe, double *conv_re,
                                                  double *conv_im, IterState *it, double reltol, double abstol,
                                                  int iter_max, double *red, int *s_last)
{
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int t = threadIdx.x, T = blockDim.x;
    __threadfence();
    __syncthreads();
    if (t == 0) {
        int prev = __hip_atomic_fetch_add(&it->counter, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        *s_last = (prev == M);
    }
    __syncthreads();
    if (!*s_last) return;
    __threadfence();
    double nre[K], nim[K];
    double mx_old = -INFINITY, mx_d = -INFINITY;
    // loads at clamped nodes first (all in flight), then the sums
    double ore[K], oim[K], vr[K][kMaxMode + 1], vi[K][kMaxMode + 1];
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int i = min(t + kk * T + 1, nr);
        ore[kk] = conv_re[i];
        oim[kk] = conv_im[i];
#pragma unroll
        for (int mm = 0; mm <= kMaxMode; mm++) {
            const int mc = min(mm, M);   // modes past M load a valid row, unused
            vr[kk][mm] = __builtin_nontemporal_load(&FX(b_spe, 1, (mc == 0 ? 0 : 2 * mc - 1), i));
            vi[kk][mm] = __builtin_nontemporal_load(&FX(b_spe, 1, 2 * mc, i));
        }
    }
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int i = t + kk * T + 1;
        nre[kk] = 0.0;
        nim[kk] = 0.0;
        if (i > nr) continue;
        mx_old = fmax(mx_old, ore[kk] * ore[kk] + oim[kk] * oim[kk]);
        double dre = ore[kk], dim = oim[kk];
#pragma unroll
        for (int mm = 0; mm <= kMaxMode; mm++) {
            if (mm > M) break;
            const double a_r = fabs(vr[kk][mm]);
            dre = dre - a_r;
            nre[kk] = nre[kk] + a_r;
            if (mm > 0) {
                const double a_i = fabs(vi[kk][mm]);
                dim = dim - a_i;
                nim[kk] = nim[kk] + a_i;
            }
        }
        mx_d = fmax(mx_d, dre * dre + dim * dim);
    }
    mx_old = block_max(mx_old, red);
    mx_d = block_max(mx_d, red);
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int i = t + kk * T + 1;
        if (i > nr) continue;
        conv_re[i] = nre[kk];
        conv_im[i] = nim[kk];
    }
    if (t == 0) {
        const double old_norm = sqrt(mx_old);
        const double ab = sqrt(mx_d);
        const double rel = (old_norm > 2.220446049250313e-16) ? ab / old_norm : 1.7976931348623157e308;
        it->rel = rel;
        it->abs_ = ab;
        it->iters = it->iters + 1;
        it->counter = 0;
        it->done = (rel < reltol || ab < abstol || it->iters >= iter_max) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// K dispatch: a solve workgroup of T threads covers K * T >= nr rows
// ---------------------------------------------------------------------------
#define QPAD_DISPATCH_K(nr, KERNEL, GRID, SHARED_EXTRA, STREAM, ARGS)                                  \
    do {                                                                                              \
        switch (rows_per_thread(nr)) {                                                                \
        case 1: hipLaunchKernelGGL(KERNEL<1>, GRID, dim3(kSolveThreads),                              \
                                   8 * stage_doubles<1>() + (SHARED_EXTRA), STREAM, ARGS); break;    \
        case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, dim3(kSolveThreads),                              \
                                   8 * stage_doubles<2>() + (SHARED_EXTRA), STREAM, ARGS); break;    \
        case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, dim3(kSolveThreads),                              \
                                   8 * stage_doubles<4>() + (SHARED_EXTRA), STREAM, ARGS); break;    \
        case 8: hipLaunchKernelGGL(KERNEL<8>, GRID, dim3(kSolveThreads),                              \
                                   8 * stage_doubles<8>() + (SHARED_EXTRA), STREAM, ARGS); break;    \
        case 16: hipLaunchKernelGGL(KERNEL<16>, GRID, dim3(kSolveThreads),                            \
                                    8 * stage_doubles<16>() + (SHARED_EXTRA), STREAM, ARGS); break;  \
        default: ::qpad::set_error("nr too large for the register-resident solve");                   \
                 return QPAD_E_INVALID;                                                               \
        }                                                                                             \
    } while (0)

constexpr int kRed = 2 * 4 * (kSolveThreads / 64);

__device__ inline FactorT dfactor(const FieldCommon &g, int kind, int m)
{
    const double *base = g.fac + (((size_t)kind * (kMaxMode + 1) + m) * 3) * g.kt;
    return FactorT{base, base + g.kt, base + 2 * g.kt};
}

// stage[r] <- v for this thread's interleaved rows (r >= nr: zero rhs)
// then blocked rows -> d
template <int K>
__device__ inline void stage_in(double *stage, const double (&v)[K], int nr, double (&d)[K])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads;
        stage[pidx<K>(r)] = (r < nr) ? v[kk] : 0.0;
    }
    __syncthreads();
    stage_to_blocked<K>(stage, d);
    __syncthreads();
}

// blocked rows d -> stage -> this thread's interleaved rows v
template <int K>
__device__ inline void stage_out(double *stage, const double (&d)[K], double (&v)[K])
{
    const int t = threadIdx.x;
    blocked_to_stage<K>(stage, d);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < K; kk++) v[kk] = stage[pidx<K>(t + kk * kSolveThreads)];
    __syncthreads();
}

}  // namespace qpad
