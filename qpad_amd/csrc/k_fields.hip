// k_fields.hip -- radial field sources, batched tridiagonal solves and the
// B_perp convergence test (FP64).  One workgroup owns one (mode, re/im)
// system: it builds the source in LDS, solves with the precomputed Thomas
// factors (block_thomas) and writes the solution with QPAD's axis rules.
#include <algorithm>
#include <cstring>

#include "kernels.h"

namespace qpad {

#define FX(p, c, k, i) (p)[((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)]

struct FieldCommon {
    int nr, ld, nm, M;
    double dr, idr, idrh, dxi, relax_fac;
    const double *fac;
};

__device__ inline Factor dfactor(const FieldCommon &g, int kind, int m)
{
    const double *base = g.fac + (((size_t)kind * (kMaxMode + 1) + m) * 3) * g.nr;
    return Factor{base, base + g.nr, base + 2 * g.nr};
}

static FieldCommon common(const qpad_ctx *c)
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
    return g;
}

__device__ inline int mode_of(int k) { return (k + 1) >> 1; }
__device__ inline bool is_im(int k) { return k > 0 && (k & 1) == 0; }
__device__ inline int partner(int k) { return k == 0 ? 0 : (is_im(k) ? k - 1 : k + 1); }

// ---------------------------------------------------------------------------
// Sources (1-based node i = 1..nr), restating the reference expressions.
// ---------------------------------------------------------------------------

// Bz source: field_b_class.f03:241-298 (own = k, part = partner(k))
__device__ double bz_source(const FieldCommon &g, const double *cu, int k, int i)
{
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int m = mode_of(k);
    const double idr = g.idr, idrh = g.idrh;
    if (m == 0) {
        if (i == 1) return -2.0 * idr * FX(cu, 1, 0, 2);
        if (i == nr) {
            double ir = idr / (double)(nr - 1);
            return -idrh * (3.0 * FX(cu, 1, 0, nr) - 4.0 * FX(cu, 1, 0, nr - 1) + FX(cu, 1, 0, nr - 2))
                   - ir * FX(cu, 1, 0, nr);
        }
        double ir = idr / (double)(i - 1);
        return -idrh * (FX(cu, 1, 0, i + 1) - FX(cu, 1, 0, i - 1)) - ir * FX(cu, 1, 0, i);
    }
    const int p = partner(k);
    const bool im = is_im(k);
    const double dm = (double)m;
    double t, ir;
    if (i == nr) {
        ir = idr / (double)(nr - 1);
        t = -idrh * (3.0 * FX(cu, 1, k, nr) - 4.0 * FX(cu, 1, k, nr - 1) + FX(cu, 1, k, nr - 2))
            - ir * FX(cu, 1, k, nr);
        return im ? t + dm * ir * FX(cu, 0, p, nr) : t - dm * ir * FX(cu, 0, p, nr);
    }
    if (i == 1) {
        if (m % 2 == 0) {
            t = -2.0 * idr * FX(cu, 1, k, 2);
            return im ? t + dm * idr * FX(cu, 0, p, 2) : t - dm * idr * FX(cu, 0, p, 2);
        }
        return 0.0;
    }
    if (i == 2 && m == 1) {
        ir = idr;
        t = -idr * (FX(cu, 1, k, 3) - FX(cu, 1, k, 2)) - ir * FX(cu, 1, k, 2);
        return im ? t + dm * ir * FX(cu, 0, p, 2) : t - dm * ir * FX(cu, 0, p, 2);
    }
    ir = idr / (double)(i - 1);
    t = -idrh * (FX(cu, 1, k, i + 1) - FX(cu, 1, k, i - 1)) - ir * FX(cu, 1, k, i);
    return im ? t + dm * ir * FX(cu, 0, p, i) : t - dm * ir * FX(cu, 0, p, i);
}

// Ez source for m > 0 (field_e_class.f03:212-268); the m = 0 source with its
// Gauss-law axis sum is handled in the kernel.
__device__ double ez_source_m(const FieldCommon &g, const double *cu, int k, int i)
{
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int m = mode_of(k), p = partner(k);
    const bool im = is_im(k);
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
    double t, ir;
    if (i == nr) {
        ir = idr / (double)(nr - 1);
        t = idrh * (3.0 * FX(cu, 0, k, nr) - 4.0 * FX(cu, 0, k, nr - 1) + FX(cu, 0, k, nr - 2))
            + ir * FX(cu, 0, k, nr);
        return im ? t + dm * ir * FX(cu, 1, p, nr) : t - dm * ir * FX(cu, 1, p, nr);
    }
    if (i == 1) {
        if (m % 2 == 0) {
            t = 2.0 * idr * FX(cu, 0, k, 2);
            return im ? t + dm * idr * FX(cu, 1, p, 2) : t - dm * idr * FX(cu, 1, p, 2);
        }
        return 0.0;
    }
    if (i == 2 && m == 1) {
        ir = idr;
        t = idr * (FX(cu, 0, k, 3) - FX(cu, 0, k, 2)) + ir * FX(cu, 0, k, 2);
        return im ? t + dm * ir * FX(cu, 1, p, 2) : t - dm * ir * FX(cu, 1, p, 2);
    }
    ir = idr / (double)(i - 1);
    t = idrh * (FX(cu, 0, k, i + 1) - FX(cu, 0, k, i - 1)) + ir * FX(cu, 0, k, i);
    return im ? t + dm * ir * FX(cu, 1, p, i) : t - dm * ir * FX(cu, 1, p, i);
}

// E_perp at node i (1..nr) of mode array k from b = bA + bB and psi
// (field_e_class.f03:443-504).  out[0] = E_r, out[1] = E_phi.
__device__ void et_node(const FieldCommon &g, const double *bA, const double *bB,
                        const double *psi, int k, int i, double out[2])
{
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int m = mode_of(k), p = partner(k);
    const double idr = g.idr, idrh = g.idrh;
    auto bv = [&](int c, int kk, int ii) {
        double v = FX(bA, c, kk, ii);
        if (bB) v = v + FX(bB, c, kk, ii);
        return v;
    };
    if (m == 0) {
        if (i == 1) { out[0] = 0.0; out[1] = 0.0; return; }
        if (i == nr)
            out[0] = bv(1, 0, nr) + idrh * (4.0 * FX(psi, 0, 0, nr - 1) - FX(psi, 0, 0, nr - 2)
                                            - 3.0 * FX(psi, 0, 0, nr));
        else
            out[0] = bv(1, 0, i) - idrh * (FX(psi, 0, 0, i + 1) - FX(psi, 0, 0, i - 1));
        out[1] = -bv(0, 0, i);
        return;
    }
    const bool im = is_im(k);
    const double dm = (double)m;
    if (i == 1) {
        if (m == 1) {
            out[0] = bv(1, k, 1) - idr * FX(psi, 0, k, 2);
            out[1] = im ? -bv(0, k, 1) - idr * FX(psi, 0, p, 2) : -bv(0, k, 1) + idr * FX(psi, 0, p, 2);
        } else {
            out[0] = 0.0;
            out[1] = 0.0;
        }
        return;
    }
    double ir = idr / (double)(i - 1);
    if (i == nr)
        out[0] = bv(1, k, nr) + idrh * (4.0 * FX(psi, 0, k, nr - 1) - FX(psi, 0, k, nr - 2)
                                        - 3.0 * FX(psi, 0, k, nr));
    else
        out[0] = bv(1, k, i) - idrh * (FX(psi, 0, k, i + 1) - FX(psi, 0, k, i - 1));
    out[1] = im ? -bv(0, k, i) - ir * dm * FX(psi, 0, p, i) : -bv(0, k, i) + ir * dm * FX(psi, 0, p, i);
}

// dJ_perp/dxi at node i of mode array k (field_src_class.f03:315-400):
// out[c] = dcu(c, k, i), c = 0 (r), 1 (phi)
__device__ void djdxi_node(const FieldCommon &g, const double *acu, const double *amu, int k,
                           int i, double out[2])
{
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int m = mode_of(k), p = partner(k);
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
    const double sg = is_im(k) ? -1.0 : 1.0;
    for (int c = 0; c < 2; c++) {
        double v;
        if (m == 0) {
            if (i == 1) v = 0.0;
            else if (i == nr) {
                double ir = idr / (double)(nr - 1);
                v = FX(acu, c, 0, nr) + idrh * (4.0 * FX(amu, c, 0, nr - 1) - FX(amu, c, 0, nr - 2)
                                                - 3.0 * FX(amu, c, 0, nr)) - ir * FX(amu, c, 0, nr);
            } else {
                double ir = idr / (double)(i - 1);
                v = FX(acu, c, 0, i) - idrh * (FX(amu, c, 0, i + 1) - FX(amu, c, 0, i - 1))
                    - ir * FX(amu, c, 0, i);
            }
        } else {
            if (i == nr) {
                double ir = idr / (double)(nr - 1);
                double t = FX(acu, c, k, nr) + idrh * (4.0 * FX(amu, c, k, nr - 1) - FX(amu, c, k, nr - 2)
                                                       - 3.0 * FX(amu, c, k, nr)) - ir * FX(amu, c, k, nr);
                v = sg > 0 ? t + dm * ir * FX(amu, c + 1, p, nr) : t - dm * ir * FX(amu, c + 1, p, nr);
            } else if (i == 1) {
                if (m == 1) {
                    double t = FX(acu, c, k, 1) - 2.0 * idr * FX(amu, c, k, 2);
                    v = sg > 0 ? t + dm * idr * FX(amu, c + 1, p, 2) : t - dm * idr * FX(amu, c + 1, p, 2);
                } else {
                    v = 0.0;
                }
            } else if (i == 2 && m == 2) {
                double ir = idr;
                double t = FX(acu, c, k, 2) - idr * (FX(amu, c, k, 3) - FX(amu, c, k, 2)) - ir * FX(amu, c, k, 2);
                v = sg > 0 ? t + dm * ir * FX(amu, c + 1, p, 2) : t - dm * ir * FX(amu, c + 1, p, 2);
            } else {
                double ir = idr / (double)(i - 1);
                double t = FX(acu, c, k, i) - idrh * (FX(amu, c, k, i + 1) - FX(amu, c, k, i - 1))
                           - ir * FX(amu, c, k, i);
                v = sg > 0 ? t + dm * ir * FX(amu, c + 1, p, i) : t - dm * ir * FX(amu, c + 1, p, i);
            }
        }
        out[c] = v;
    }
}

// ---------------------------------------------------------------------------
// Block reductions
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *red)
{
    // fixed butterfly within waves, then waves in order: deterministic
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

// ---------------------------------------------------------------------------
// psi (+ q combine), Bz from cu, convergence 'record'
// ---------------------------------------------------------------------------
struct PsiBzArgs {
    FieldCommon g;
    double *q_spe, *psi, *b_spe;
    const double *cu;
    double *conv_re, *conv_im;
    IterState *it;
    int nspecies;
    const double *P0[kMaxSpecies], *P1[kMaxSpecies], *qn[kMaxSpecies];
    int do_psi, do_bz, do_record;
    int from_partials;   // 1: assemble q_spe from the deposit partials first
};

__global__ __launch_bounds__(kSolveThreads) void k_psi_bz(PsiBzArgs a)
{
    extern __shared__ double smem[];
    __shared__ double red[32];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int T = blockDim.x, t = threadIdx.x;
    int b = blockIdx.x;
    double *rhs = smem;
    if (b < nm) {
        if (!a.do_psi) return;
        const int k = b, m = mode_of(k);
        for (int i = t; i <= nr + 1; i += T) {
            if (!a.from_partials) {
                if (i >= 1 && i <= nr) rhs[i - 1] = -1.0 * a.q_spe[(size_t)k * ld + i];
                continue;
            }
            double v = 0.0;
            for (int s = 0; s < a.nspecies; s++) {
                if (i >= 1 && i <= nr) v += a.P0[s][(size_t)(i - 1) * nm + k];
                if (i >= 2) v += a.P1[s][(size_t)(i - 2) * nm + k];
            }
            // axis rules and 1/r of qdeposit_part2d (part2d_class.f03:313-336)
            if (i == 0) v = 0.0;
            else if (i == 1) v = (k == 0) ? 8.0 * v : 0.0;
            else v = v * (1.0 / (double)(i - 1));
            for (int s = 0; s < a.nspecies; s++)
                if (a.qn[s]) v = v + a.qn[s][(size_t)k * ld + i];
            a.q_spe[(size_t)k * ld + i] = v;
            if (i >= 1 && i <= nr) rhs[i - 1] = -1.0 * v;
        }
        __syncthreads();
        block_thomas(dfactor(g, FK_PSI, m), rhs, nr, red);
        for (int i = t + 1; i <= nr; i += T) {
            double v = rhs[i - 1];
            if (m > 0 && i == 1) v = 0.0;
            a.psi[(size_t)k * ld + i] = v;
        }
        return;
    }
    b -= nm;
    if (b < nm) {
        if (!a.do_bz) return;
        const int k = b, m = mode_of(k);
        for (int i = t + 1; i <= nr; i += T) rhs[i - 1] = bz_source(g, a.cu, k, i);
        __syncthreads();
        block_thomas(dfactor(g, FK_BZ, m), rhs, nr, red);
        for (int i = t + 1; i <= nr; i += T) {
            double v = rhs[i - 1];
            if (m > 0 && i == 1) v = 0.0;
            FX(a.b_spe, 2, k, i) = v;
        }
        return;
    }
    // convergence_tester 'record' (simulation_class.f03:549-558)
    if (!a.do_record) return;
    if (t == 0) {
        a.it->done = 0;
        a.it->iters = 0;
        a.it->counter = 0;
    }
    for (int i = t + 1; i <= nr; i += T) {
        double re = 0.0, im = 0.0;
        for (int m = 0; m <= g.M; m++) {
            re = re + fabs(FX(a.b_spe, 1, kre(m), i));
            if (m > 0) im = im + fabs(FX(a.b_spe, 1, kim(m), i));
        }
        a.conv_re[i] = re;
        a.conv_im[i] = im;
    }
}

int launch_psi_bz(qpad_ctx *c, bool do_psi, bool do_bz, bool do_record)
{
    PsiBzArgs a;
    a.g = common(c);
    a.q_spe = c->f[QPAD_F_Q_SPE];
    a.psi = c->f[QPAD_F_PSI];
    a.b_spe = c->f[QPAD_F_B_SPE];
    a.cu = c->f[QPAD_F_CU];
    a.conv_re = c->conv_re;
    a.conv_im = c->conv_im;
    a.it = c->it;
    a.nspecies = c->nspecies;
    for (int s = 0; s < kMaxSpecies; s++) {
        a.P0[s] = s < c->nspecies ? c->sp[s].P0 : nullptr;
        a.P1[s] = s < c->nspecies ? c->sp[s].P1 : nullptr;
        a.qn[s] = (s < c->nspecies && c->sp[s].cfg.neutralized) ? c->sp[s].qn : nullptr;
    }
    a.do_psi = do_psi;
    a.do_bz = do_bz;
    a.do_record = do_record;
    a.from_partials = 1;
    int nblk = 2 * c->g.nm + 1;
    size_t sh = sizeof(double) * c->g.nr;
    hipLaunchKernelGGL(k_psi_bz, dim3(nblk), dim3(kSolveThreads), sh, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_psi_from_qspe(qpad_ctx *c)
{
    PsiBzArgs a;
    memset(&a, 0, sizeof(a));
    a.g = common(c);
    a.q_spe = c->f[QPAD_F_Q_SPE];
    a.psi = c->f[QPAD_F_PSI];
    a.do_psi = 1;
    size_t sh = sizeof(double) * c->g.nr;
    hipLaunchKernelGGL(k_psi_bz, dim3(c->g.nm), dim3(kSolveThreads), sh, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Ez + E_perp (+ b = b_spe + b_beam, + slice epilogue)
// ---------------------------------------------------------------------------
struct EzEtArgs {
    FieldCommon g;
    const double *jay;      // Ez source (nullptr: no Ez)
    double *ez_dst;
    const double *bA, *bB;  // E_perp from bA + bB (bB may be nullptr)
    double *b_out;          // write bA + bB here (nullptr: no)
    double *et_dst;         // nullptr: no E_perp
    const double *psi;
    double *et2_dst;        // second E_perp from bA alone (e_spe), nullptr: no
    IterState *it;
    int check_done;
    // slice epilogue (simulation_class.f03:403-456)
    int final;
    double *q_spe;
    const double *cu, *b_spe, *q_beam_j;
    double *h_e, *h_b, *h_psi, *h_b_spe, *h_e_spe, *h_cu, *h_q_spe;
    double *q_beam_f1, *b_beam_f1;
};

__global__ __launch_bounds__(kSolveThreads) void k_ez_et(EzEtArgs a)
{
    extern __shared__ double smem[];
    __shared__ double red[32];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int T = blockDim.x, t = threadIdx.x;
    const int k = blockIdx.x, m = mode_of(k);
    if (a.check_done && a.it->done) return;
    double *rhs = smem;

    // elementwise part: b, E_perp (+ epilogue copies)
    for (int i = t; i <= nr + 1; i += T) {
        if (a.b_out) {
            for (int cc = 0; cc < 3; cc++) {
                double v = FX(a.bA, cc, k, i);
                if (a.bB) v = v + FX(a.bB, cc, k, i);
                FX(a.b_out, cc, k, i) = v;
                if (a.final) FX(a.h_b, cc, k, i) = v;
            }
        }
        if (a.et_dst) {
            double ev[2];
            if (i >= 1 && i <= nr) {
                et_node(g, a.bA, a.bB, a.psi, k, i, ev);
                FX(a.et_dst, 0, k, i) = ev[0];
                FX(a.et_dst, 1, k, i) = ev[1];
            } else {
                ev[0] = FX(a.et_dst, 0, k, i);
                ev[1] = FX(a.et_dst, 1, k, i);
            }
            if (a.final) {
                FX(a.h_e, 0, k, i) = ev[0];
                FX(a.h_e, 1, k, i) = ev[1];
            }
        }
        if (a.et2_dst) {
            double ev[2];
            if (i >= 1 && i <= nr) {
                et_node(g, a.bA, nullptr, a.psi, k, i, ev);
                FX(a.et2_dst, 0, k, i) = ev[0];
                FX(a.et2_dst, 1, k, i) = ev[1];
            } else {
                ev[0] = FX(a.et2_dst, 0, k, i);
                ev[1] = FX(a.et2_dst, 1, k, i);
            }
            if (a.final) {
                FX(a.h_e_spe, 0, k, i) = ev[0];
                FX(a.h_e_spe, 1, k, i) = ev[1];
                FX(a.h_e_spe, 2, k, i) = FX(a.et2_dst, 2, k, i);
            }
        }
        if (a.final) {
            // cbq / cu -> f2 / q_spe += Jz / q_spe -> f2 (simulation_class.f03:403-411)
            for (int cc = 0; cc < 3; cc++) FX(a.h_cu, cc, k, i) = FX(a.cu, cc, k, i);
            double q = a.q_spe[(size_t)k * ld + i] + FX(a.cu, 2, k, i);
            a.q_spe[(size_t)k * ld + i] = q;
            a.h_q_spe[(size_t)k * ld + i] = q;
            a.h_psi[(size_t)k * ld + i] = a.psi[(size_t)k * ld + i];
            for (int cc = 0; cc < 3; cc++) FX(a.h_b_spe, cc, k, i) = FX(a.b_spe, cc, k, i);
            a.q_beam_f1[(size_t)k * ld + i] = a.q_beam_j[(size_t)k * ld + i];
            for (int cc = 0; cc < 3; cc++) FX(a.b_beam_f1, cc, k, i) = a.bB ? FX(a.bB, cc, k, i) : 0.0;
            if (i == 0 || i == nr + 1) FX(a.h_e, 2, k, i) = FX(a.ez_dst, 2, k, i);
        }
    }
    if (!a.jay) return;

    // Ez source (field_e_class.f03:182-275) and solve
    if (m == 0) {
        double part = 0.0;
        for (int i = t + 1; i <= nr; i += T) {
            const double *cu = a.jay;
            double v;
            double idr = g.idr, idrh = g.idrh;
            if (i >= 2 && i <= nr - 2) {
                double ir = idr / (double)(i - 1);
                v = idrh * (FX(cu, 0, 0, i + 1) - FX(cu, 0, 0, i - 1)) + ir * FX(cu, 0, 0, i);
                part += v * (double)(i - 1);
            } else if (i == nr - 1) {
                double ir = idr / (double)(nr - 2);
                v = idrh * (FX(cu, 0, 0, nr) - FX(cu, 0, 0, nr - 2)) + ir * FX(cu, 0, 0, nr - 1);
            } else if (i == nr) {
                double ir = idr / (double)(nr - 1);
                v = idr * (FX(cu, 0, 0, nr) - FX(cu, 0, 0, nr - 1)) + ir * FX(cu, 0, 0, nr);
            } else {
                v = 0.0;  // i == 1, set below
            }
            rhs[i - 1] = v;
        }
        double div = block_sum(part, red);
        if (t == 0) {
            const double *cu = a.jay;
            div = div - g.idrh * (FX(cu, 0, 0, nr - 2) + FX(cu, 0, 0, nr - 1)) * ((double)nr - 2.5);
            rhs[0] = -8.0 * div;
        }
    } else {
        for (int i = t + 1; i <= nr; i += T) rhs[i - 1] = ez_source_m(g, a.jay, k, i);
    }
    __syncthreads();
    block_thomas(dfactor(g, FK_EZ, m), rhs, nr, red);
    for (int i = t + 1; i <= nr; i += T) {
        double v = rhs[i - 1];
        if (m > 0 && i == 1) v = 0.0;
        FX(a.ez_dst, 2, k, i) = v;
        if (a.final) FX(a.h_e, 2, k, i) = v;
    }
}

static EzEtArgs ezet_base(qpad_ctx *c)
{
    EzEtArgs a;
    memset(&a, 0, sizeof(a));
    a.g = common(c);
    a.psi = c->f[QPAD_F_PSI];
    a.it = c->it;
    return a;
}

static int run_ezet(qpad_ctx *c, const EzEtArgs &a)
{
    size_t sh = sizeof(double) * c->g.nr;
    hipLaunchKernelGGL(k_ez_et, dim3(c->g.nm), dim3(kSolveThreads), sh, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

static inline size_t slice_off(const qpad_ctx *c, int ncomp, int j)
{
    return (size_t)j * ncomp * c->g.nm * c->g.ld;
}

int launch_ez_et(qpad_ctx *c, int j, bool final, bool check_done)
{
    EzEtArgs a = ezet_base(c);
    a.jay = c->f[QPAD_F_CU];
    a.ez_dst = c->f[QPAD_F_E];
    a.bA = c->f[QPAD_F_B_SPE];
    a.bB = c->h_b_beam + slice_off(c, 3, j);
    a.b_out = c->f[QPAD_F_B];
    a.et_dst = c->f[QPAD_F_E];
    a.check_done = check_done;
    if (final) {
        a.final = 1;
        a.et2_dst = c->f[QPAD_F_E_SPE];
        a.q_spe = c->f[QPAD_F_Q_SPE];
        a.cu = c->f[QPAD_F_CU];
        a.b_spe = c->f[QPAD_F_B_SPE];
        a.q_beam_j = c->h_q_beam + slice_off(c, 1, j);
        a.h_e = c->h_e + slice_off(c, 3, j);
        a.h_b = c->h_b + slice_off(c, 3, j);
        a.h_psi = c->h_psi + slice_off(c, 1, j);
        a.h_b_spe = c->h_b_spe + slice_off(c, 3, j);
        a.h_e_spe = c->h_e_spe + slice_off(c, 3, j);
        a.h_cu = c->h_cu + slice_off(c, 3, j);
        a.h_q_spe = c->h_q_spe + slice_off(c, 1, j);
        a.q_beam_f1 = c->f[QPAD_F_Q_BEAM];
        a.b_beam_f1 = c->f[QPAD_F_B_BEAM];
    }
    return run_ezet(c, a);
}

int launch_ez_only(qpad_ctx *c, int dst, int jay)
{
    EzEtArgs a = ezet_base(c);
    a.jay = c->f[jay];
    a.ez_dst = c->f[dst];
    return run_ezet(c, a);
}

int launch_et_only(qpad_ctx *c, int dst, int bsrc)
{
    EzEtArgs a = ezet_base(c);
    a.bA = c->f[bsrc];
    a.et_dst = c->f[dst];
    return run_ezet(c, a);
}

int launch_add_b(qpad_ctx *c, int j)
{
    EzEtArgs a = ezet_base(c);
    a.bA = c->f[QPAD_F_B_SPE];
    a.bB = c->f[QPAD_F_B_BEAM];
    a.b_out = c->f[QPAD_F_B];
    (void)j;
    return run_ezet(c, a);
}

// ---------------------------------------------------------------------------
// cu, acu, amu from the per-cell partials (axis rules of
// amjdeposit_robust_part2d, part2d_class.f03:914-1004)
// ---------------------------------------------------------------------------
struct CombineJArgs {
    int nr, ld, nm;
    double *cu, *acu, *amu;
    int nspecies;
    const double *P0[kMaxSpecies], *P1[kMaxSpecies];
    const IterState *it;
    int check_done;
    int accumulate;
};

__global__ __launch_bounds__(256) void k_combine_j(CombineJArgs a)
{
    if (a.check_done && a.it->done) return;
    const int nm = a.nm, ld = a.ld, nr = a.nr;
    const int NV = 8 * nm;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)NV * (nr + 2);
    if (tid >= total) return;
    const int i = (int)(tid % (nr + 2));
    const int v = (int)(tid / (nr + 2));
    const int comp = v / nm, k = v % nm, m = mode_of(k);
    double val = 0.0;
    for (int s = 0; s < a.nspecies; s++) {
        if (i >= 1 && i <= nr) val += a.P0[s][(size_t)(i - 1) * NV + v];
        if (i >= 2) val += a.P1[s][(size_t)(i - 2) * NV + v];
    }
    if (i == 0) {
        val = 0.0;
    } else if (i == 1) {
        // comp 0..2 = cu (r, phi, z), 3..4 = acu (r, phi), 5..7 = amu
        double f;
        if (m == 0) f = (comp == 2) ? 8.0 : 0.0;
        else if (m == 1) f = (comp <= 1 || comp == 3 || comp == 4) ? 8.0 : 0.0;
        else if (m == 2) f = (comp >= 5) ? 8.0 : 0.0;
        else f = 0.0;
        val = (f == 0.0) ? 0.0 : f * val;
    } else {
        val = val * (1.0 / (double)(i - 1));
    }
    double *dst = (comp < 3) ? &FX(a.cu, comp, k, i)
                  : (comp < 5) ? &FX(a.acu, comp - 3, k, i) : &FX(a.amu, comp - 5, k, i);
    *dst = a.accumulate ? *dst + val : val;
}

int launch_combine_j(qpad_ctx *c, bool check_done)
{
    CombineJArgs a;
    a.nr = c->g.nr;
    a.ld = c->g.ld;
    a.nm = c->g.nm;
    a.cu = c->f[QPAD_F_CU];
    a.acu = c->f[QPAD_F_ACU];
    a.amu = c->f[QPAD_F_AMU];
    a.nspecies = c->nspecies;
    for (int s = 0; s < kMaxSpecies; s++) {
        a.P0[s] = s < c->nspecies ? c->sp[s].P0 : nullptr;
        a.P1[s] = s < c->nspecies ? c->sp[s].P1 : nullptr;
    }
    a.it = c->it;
    a.check_done = check_done;
    a.accumulate = 0;
    int64_t total = (int64_t)8 * c->g.nm * (c->g.nr + 2);
    int nblk = (int)((total + 255) / 256);
    hipLaunchKernelGGL(k_combine_j, dim3(nblk), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_combine_j_species(qpad_ctx *c, int sp)
{
    CombineJArgs a;
    memset(&a, 0, sizeof(a));
    a.nr = c->g.nr;
    a.ld = c->g.ld;
    a.nm = c->g.nm;
    a.cu = c->f[QPAD_F_CU];
    a.acu = c->f[QPAD_F_ACU];
    a.amu = c->f[QPAD_F_AMU];
    a.nspecies = 1;
    a.P0[0] = c->sp[sp].P0;
    a.P1[0] = c->sp[sp].P1;
    a.it = c->it;
    a.accumulate = 1;
    int64_t total = (int64_t)8 * c->g.nm * (c->g.nr + 2);
    hipLaunchKernelGGL(k_combine_j, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// dJ/dxi + plasma B_perp (B+/B- Helmholtz) + Bz + convergence 'compare'
// ---------------------------------------------------------------------------
struct BperpArgs {
    FieldCommon g;
    const double *acu, *amu, *cu;
    double *dcu, *b_spe;
    double *conv_re, *conv_im;
    IterState *it;
    double reltol, abstol;
    int iter_max;
    int do_bperp, do_bz, do_compare, check_done, do_djdxi_only;
    int compare_only;   // skip dJ/dxi and the solves, run 'compare' only
};

__global__ __launch_bounds__(kSolveThreads) void k_bperp_bz(BperpArgs a)
{
    extern __shared__ double smem[];
    __shared__ double red[32];
    __shared__ int s_last;
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int T = blockDim.x, t = threadIdx.x;
    if (a.check_done && a.it->done) return;
    int b = blockIdx.x;
    if (b <= M && a.compare_only) goto compare;
    if (b <= M) {
        if (!a.do_bperp && !a.do_djdxi_only) return;
        const int m = b, kr = kre(m), ki = kim(m);
        double *b1r = smem, *b2r = smem + nr, *b1i = smem + 2 * nr, *b2i = smem + 3 * nr;
        const double idr = g.idr, idrh = g.idrh, dm = (double)m;
        const double rx = g.relax_fac * (idr * idr);
        const double *f2 = a.cu, *f3 = a.b_spe;
        for (int i = t + 1; i <= nr; i += T) {
            double dr_[2], di_[2] = {0.0, 0.0};
            djdxi_node(g, a.acu, a.amu, kr, i, dr_);
            FX(a.dcu, 0, kr, i) = dr_[0];
            FX(a.dcu, 1, kr, i) = dr_[1];
            if (m > 0) {
                djdxi_node(g, a.acu, a.amu, ki, i, di_);
                FX(a.dcu, 0, ki, i) = di_[0];
                FX(a.dcu, 1, ki, i) = di_[1];
            }
            if (a.do_djdxi_only) continue;
            // B+/B- sources (field_b_class.f03:388-497)
            if (m == 0) {
                double s1, s2;
                if (i == 1) { s1 = 0.0; s2 = 0.0; }
                else if (i == nr) {
                    s1 = -dr_[1] - FX(f3, 0, 0, nr) * rx;
                    s2 = dr_[0] + idrh * (3.0 * FX(f2, 2, 0, nr) - 4.0 * FX(f2, 2, 0, nr - 1) + FX(f2, 2, 0, nr - 2))
                         - FX(f3, 1, 0, nr) * rx;
                } else if (i == 2) {
                    s1 = -dr_[1] - FX(f3, 0, 0, i) * rx;
                    s2 = dr_[0] + idr * (FX(f2, 2, 0, 3) - FX(f2, 2, 0, 2)) - FX(f3, 1, 0, 2) * rx;
                } else {
                    s1 = -dr_[1] - FX(f3, 0, 0, i) * rx;
                    s2 = dr_[0] + idrh * (FX(f2, 2, 0, i + 1) - FX(f2, 2, 0, i - 1)) - FX(f3, 1, 0, i) * rx;
                }
                b1r[i - 1] = s1;
                b2r[i - 1] = s2;
            } else {
                double s1r, s1i, s2r, s2i;
                if (i == 1) {
                    if (m == 1) {
                        s1r = -dr_[1] + idr * dm * FX(f2, 2, ki, 2);
                        s1i = -di_[1] - idr * dm * FX(f2, 2, kr, 2);
                        s2r = dr_[0] + idr * FX(f2, 2, kr, 2);
                        s2i = di_[0] + idr * FX(f2, 2, ki, 2);
                    } else if (m % 2 == 0) {
                        s1r = 0.0; s1i = 0.0; s2r = 0.0; s2i = 0.0;
                    } else {
                        s1r = idr * dm * FX(f2, 2, ki, 2);
                        s1i = idr * dm * FX(f2, 2, kr, 2);
                        s2r = idr * FX(f2, 2, kr, 2);
                        s2i = idr * FX(f2, 2, ki, 2);
                    }
                } else if (i == nr) {
                    double ir = idr / (double)(nr - 1);
                    s1r = -dr_[1] + dm * FX(f2, 2, ki, nr) * ir;
                    s1i = -di_[1] - dm * FX(f2, 2, kr, nr) * ir;
                    s2r = dr_[0] + idrh * (3.0 * FX(f2, 2, kr, nr) - 4.0 * FX(f2, 2, kr, nr - 1) + FX(f2, 2, kr, nr - 2));
                    s2i = di_[0] + idrh * (3.0 * FX(f2, 2, ki, nr) - 4.0 * FX(f2, 2, ki, nr - 1) + FX(f2, 2, ki, nr - 2));
                } else {
                    double ir = idr / (double)(i - 1);
                    s1r = -dr_[1] + dm * FX(f2, 2, ki, i) * ir;
                    s1i = -di_[1] - dm * FX(f2, 2, kr, i) * ir;
                    s2r = dr_[0] + idrh * (FX(f2, 2, kr, i + 1) - FX(f2, 2, kr, i - 1));
                    s2i = di_[0] + idrh * (FX(f2, 2, ki, i + 1) - FX(f2, 2, ki, i - 1));
                }
                const double f3r0 = FX(f3, 0, kr, i), f3r1 = FX(f3, 1, kr, i);
                const double f3i0 = FX(f3, 0, ki, i), f3i1 = FX(f3, 1, ki, i);
                b1r[i - 1] = s1r - s2i - (f3r0 - f3i1) * rx;
                b1i[i - 1] = s1i + s2r - (f3i0 + f3r1) * rx;
                b2r[i - 1] = s1r + s2i - (f3r0 + f3i1) * rx;
                b2i[i - 1] = s1i - s2r - (f3i0 - f3r1) * rx;
            }
        }
        if (a.do_djdxi_only) return;
        __syncthreads();
        block_thomas(dfactor(g, FK_BPLUS, m), b1r, nr, red);
        block_thomas(dfactor(g, FK_BMINUS, m), b2r, nr, red);
        if (m > 0) {
            block_thomas(dfactor(g, FK_BPLUS, m), b1i, nr, red);
            block_thomas(dfactor(g, FK_BMINUS, m), b2i, nr, red);
        }
        // recombination (field_b_class.f03:715-752)
        for (int i = t + 1; i <= nr; i += T) {
            if (m == 0) {
                double br = b1r[i - 1], bp = b2r[i - 1];
                if (i == 1) { br = 0.0; bp = 0.0; }
                FX(a.b_spe, 0, 0, i) = br;
                FX(a.b_spe, 1, 0, i) = bp;
            } else {
                double v0r = 0.5 * (b1r[i - 1] + b2r[i - 1]);
                double v0i = 0.5 * (b1i[i - 1] + b2i[i - 1]);
                double v1r = 0.5 * (b1i[i - 1] - b2i[i - 1]);
                double v1i = 0.5 * (-b1r[i - 1] + b2r[i - 1]);
                if (i == 1 && m != 1) { v0r = 0.0; v0i = 0.0; v1r = 0.0; v1i = 0.0; }
                FX(a.b_spe, 0, kr, i) = v0r;
                FX(a.b_spe, 0, ki, i) = v0i;
                FX(a.b_spe, 1, kr, i) = v1r;
                FX(a.b_spe, 1, ki, i) = v1i;
            }
        }
        if (!a.do_compare) return;
    }
    if (b <= M) {
    compare:
        // last B_perp block to finish runs convergence_tester 'compare'
        __threadfence();
        __syncthreads();
        if (t == 0) {
            int prev = __hip_atomic_fetch_add(&a.it->counter, 1, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
            s_last = (prev == M);
        }
        __syncthreads();
        if (!s_last) return;
        __threadfence();
        double mx_old = -INFINITY, mx_d = -INFINITY;
        for (int i = t + 1; i <= nr; i += T) {
            double ore = a.conv_re[i], oim = a.conv_im[i];
            mx_old = fmax(mx_old, ore * ore + oim * oim);
            double dre = ore, dim = oim, nre = 0.0, nim = 0.0;
            for (int mm = 0; mm <= M; mm++) {
                double vr = fabs(__builtin_nontemporal_load(&FX(a.b_spe, 1, kre(mm), i)));
                dre = dre - vr;
                nre = nre + vr;
                if (mm > 0) {
                    double vi = fabs(__builtin_nontemporal_load(&FX(a.b_spe, 1, kim(mm), i)));
                    dim = dim - vi;
                    nim = nim + vi;
                }
            }
            mx_d = fmax(mx_d, dre * dre + dim * dim);
            a.conv_re[i] = nre;
            a.conv_im[i] = nim;
        }
        mx_old = block_max(mx_old, red);
        mx_d = block_max(mx_d, red);
        if (t == 0) {
            double old_norm = sqrt(mx_old);
            double ab = sqrt(mx_d);
            double rel = (old_norm > 2.220446049250313e-16) ? ab / old_norm : 1.7976931348623157e308;
            IterState *it = a.it;
            it->rel = rel;
            it->abs_ = ab;
            it->iters = it->iters + 1;
            it->counter = 0;
            it->done = (rel < a.reltol || ab < a.abstol || it->iters >= a.iter_max) ? 1 : 0;
        }
        return;
    }
    b -= (M + 1);
    if (b < nm) {
        if (!a.do_bz) return;
        const int k = b, m = mode_of(k);
        double *rhs = smem;
        for (int i = t + 1; i <= nr; i += T) rhs[i - 1] = bz_source(g, a.cu, k, i);
        __syncthreads();
        block_thomas(dfactor(g, FK_BZ, m), rhs, nr, red);
        for (int i = t + 1; i <= nr; i += T) {
            double v = rhs[i - 1];
            if (m > 0 && i == 1) v = 0.0;
            FX(a.b_spe, 2, k, i) = v;
        }
    }
}

static BperpArgs bperp_base(qpad_ctx *c)
{
    BperpArgs a;
    memset(&a, 0, sizeof(a));
    a.g = common(c);
    a.acu = c->f[QPAD_F_ACU];
    a.amu = c->f[QPAD_F_AMU];
    a.cu = c->f[QPAD_F_CU];
    a.dcu = c->f[QPAD_F_DCU];
    a.b_spe = c->f[QPAD_F_B_SPE];
    a.conv_re = c->conv_re;
    a.conv_im = c->conv_im;
    a.it = c->it;
    a.reltol = c->g.reltol;
    a.abstol = c->g.abstol;
    a.iter_max = c->g.iter_max;
    return a;
}

int launch_compare_only(qpad_ctx *c);

static int run_bperp(qpad_ctx *c, const BperpArgs &a)
{
    int nblk = (c->g.M + 1) + c->g.nm;
    size_t sh = sizeof(double) * 4 * c->g.nr;
    hipLaunchKernelGGL(k_bperp_bz, dim3(nblk), dim3(kSolveThreads), sh, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_bperp_bz(qpad_ctx *c, bool do_bperp, bool do_bz, bool do_compare, bool check_done)
{
    BperpArgs a = bperp_base(c);
    a.do_bperp = do_bperp;
    a.do_bz = do_bz;
    a.do_compare = do_compare;
    a.check_done = check_done;
    return run_bperp(c, a);
}

int launch_compare_only(qpad_ctx *c)
{
    BperpArgs a = bperp_base(c);
    a.compare_only = 1;
    a.do_compare = 1;
    return run_bperp(c, a);
}

int launch_djdxi_only(qpad_ctx *c)
{
    BperpArgs a = bperp_base(c);
    a.do_djdxi_only = 1;
    return run_bperp(c, a);
}

int launch_bz_only(qpad_ctx *c, int dst, int jay)
{
    PsiBzArgs a;
    memset(&a, 0, sizeof(a));
    a.g = common(c);
    a.cu = c->f[jay];
    a.b_spe = c->f[dst];
    a.do_bz = 1;
    int nblk = 2 * c->g.nm + 1;
    size_t sh = sizeof(double) * c->g.nr;
    hipLaunchKernelGGL(k_psi_bz, dim3(nblk), dim3(kSolveThreads), sh, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Beam B_perp for every slice (field_b_class.f03:309-358, :545-701)
// ---------------------------------------------------------------------------
struct BtBeamArgs {
    FieldCommon g;
    const double *hq;   // first q_beam slice to solve ([1][nm][ld] per slice)
    double *hb;         // first b_beam slice to write ([3][nm][ld] per slice)
    size_t qstride, bstride;
};

__global__ __launch_bounds__(kSolveThreads) void k_bt_beam_all(BtBeamArgs a)
{
    extern __shared__ double smem[];
    __shared__ double red[32];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int T = blockDim.x, t = threadIdx.x;
    const int m = blockIdx.y;
    const int kr = kre(m), ki = kim(m);
    const double *q = a.hq + (size_t)blockIdx.x * a.qstride;
    double *bf = a.hb + (size_t)blockIdx.x * a.bstride;
    double *bre = smem, *bim = smem + nr;
    for (int i = t + 1; i <= nr; i += T) {
        bre[i - 1] = -1.0 * q[(size_t)kr * ld + i];
        if (m > 0) bim[i - 1] = -1.0 * q[(size_t)ki * ld + i];
    }
    __syncthreads();
    block_thomas(dfactor(g, FK_BT, m), bre, nr, red);
    if (m > 0) block_thomas(dfactor(g, FK_BT, m), bim, nr, red);
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
    // buffers are 1-based in the reference: buf(i) = bre[i-1]
    for (int i = t + 1; i <= nr; i += T) {
        if (m == 0) {
            double v0 = 0.0, v1;
            if (i == 1) v1 = 0.0;
            else if (i == nr) v1 = -idrh * (3.0 * bre[nr - 1] - 4.0 * bre[nr - 2] + bre[nr - 3]);
            else v1 = -idrh * (bre[i] - bre[i - 2]);
            FX(bf, 0, 0, i) = v0;
            FX(bf, 1, 0, i) = v1;
        } else {
            double r0, i0, r1, i1;
            if (i == 1) {
                if (m == 1) {
                    r0 = -idr * dm * bim[1];
                    i0 = idr * dm * bre[1];
                    r1 = -idr * bre[1];
                    i1 = -idr * bim[1];
                } else {
                    r0 = 0.0; i0 = 0.0; r1 = 0.0; i1 = 0.0;
                }
            } else if (i == nr) {
                double ir = idr / (double)(nr - 1);
                r0 = -ir * dm * bim[nr - 1];
                i0 = ir * dm * bre[nr - 1];
                r1 = -idrh * (3.0 * bre[nr - 1] - 4.0 * bre[nr - 2] + bre[nr - 3]);
                i1 = -idrh * (3.0 * bim[nr - 1] - 4.0 * bim[nr - 2] + bim[nr - 3]);
            } else {
                double ir = idr / (double)(i - 1);
                r0 = -ir * dm * bim[i - 1];
                i0 = ir * dm * bre[i - 1];
                r1 = -idrh * (bre[i] - bre[i - 2]);
                i1 = -idrh * (bim[i] - bim[i - 2]);
            }
            FX(bf, 0, kr, i) = r0;
            FX(bf, 0, ki, i) = i0;
            FX(bf, 1, kr, i) = r1;
            FX(bf, 1, ki, i) = i1;
        }
    }
}

int launch_bt_beam_all(qpad_ctx *c)
{
    BtBeamArgs a;
    a.g = common(c);
    a.qstride = (size_t)c->g.nm * c->g.ld;
    a.bstride = (size_t)3 * c->g.nm * c->g.ld;
    a.hq = c->h_q_beam + a.qstride;   // slices 1..nxi
    a.hb = c->h_b_beam + a.bstride;
    size_t sh = sizeof(double) * 2 * c->g.nr;
    hipLaunchKernelGGL(k_bt_beam_all, dim3(c->g.nxi, c->g.M + 1), dim3(kSolveThreads), sh,
                       c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_bt_beam_one(qpad_ctx *c, int j)
{
    BtBeamArgs a;
    a.g = common(c);
    a.qstride = 0;
    a.bstride = 0;
    a.hq = c->h_q_beam + (size_t)j * c->g.nm * c->g.ld;
    a.hb = c->f[QPAD_F_B_BEAM];
    // q_beam%copy_slice(j, p_copy_2to1)
    QPAD_HIP_CHECK(hipMemcpyAsync(c->f[QPAD_F_Q_BEAM], a.hq, sizeof(double) * c->g.nm * c->g.ld,
                                  hipMemcpyDeviceToDevice, c->stream));
    size_t sh = sizeof(double) * 2 * c->g.nr;
    hipLaunchKernelGGL(k_bt_beam_all, dim3(1, c->g.M + 1), dim3(kSolveThreads), sh, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// allow the LDS-resident solves to use more than the default dynamic LDS
int field_kernels_init(size_t dyn_lds)
{
    const void *ks[] = {(const void *)k_psi_bz, (const void *)k_ez_et, (const void *)k_bperp_bz,
                        (const void *)k_bt_beam_all};
    for (const void *k : ks)
        QPAD_HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds));
    return 0;
}

// ---------------------------------------------------------------------------
// J_perp predictor for the next slice (simulation_class.f03:425-426)
// ---------------------------------------------------------------------------
__global__ void k_predictor(double *cu, double *dcu, double dxi, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double d = dxi * dcu[t];
    dcu[t] = d;
    cu[t] = cu[t] + d;
}

int launch_predictor(qpad_ctx *c)
{
    int64_t n = (int64_t)2 * c->g.nm * c->g.ld;
    hipLaunchKernelGGL(k_predictor, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->f[QPAD_F_CU], c->f[QPAD_F_DCU], c->g.dxi, n);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// generic fill / copy (used instead of hipMemsetAsync so that captured
// graphs never contain memset nodes)
// ---------------------------------------------------------------------------
__global__ void k_fill(double *p, int64_t n, double v)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) p[t] = v;
}

__global__ void k_copy(double *d, const double *s, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) d[t] = s[t];
}

int launch_fill(qpad_ctx *c, double *p, size_t n, double v)
{
    if (n == 0) return 0;
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, c->stream, p, (int64_t)n, v);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_copy(qpad_ctx *c, double *dst, const double *src, size_t n)
{
    if (n == 0) return 0;
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c->stream, dst, src, (int64_t)n);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace qpad
