// k_fields.hip -- the per-call field solves behind the fine-grained entry
// points (field_psi%solve, field_e%solve, field_b%solve, field_djdxi%solve,
// convergence_tester), the psi/Bz stage of each slice, the first E of each
// slice and the beam B of every slice (FP64).  One workgroup owns one
// (mode, re/im) system: coalesced global I/O, a padded LDS stage into the
// blocked row order, and the register-resident tridiagonal solve of
// tridiag.h with the precomputed Thomas factors.
#include <algorithm>
#include <cstring>

#include "field_src.h"
#include "kernels.h"
#include "tridiag.h"

namespace qpad {


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
    const double *O0[kMaxSpecies], *O1[kMaxSpecies];
    const int32_t *os[kMaxSpecies];
    int do_psi, do_bz, do_record;
    int from_partials;   // 1: assemble q_spe from the deposit partials first
};

// rho - Jz of every species at node i of mode array k (axis rules and 1/r of
// qdeposit_part2d, part2d_class.f03:313-336) plus the backgrounds
__device__ inline double q_node(const PsiBzArgs &a, int k, int i)
{
    const int nm = a.g.nm, ld = a.g.ld, nr = a.g.nr;
    double v = 0.0;
    for (int s = 0; s < a.nspecies; s++) {
        if (i >= 1 && i <= nr) v += cell_partial(a.P0[s], a.O0[s], a.os[s], i - 1, nm, k);
        if (i >= 2) v += cell_partial(a.P1[s], a.O1[s], a.os[s], i - 2, nm, k);
    }
    if (i == 0) v = 0.0;
    else if (i == 1) v = (k == 0) ? 8.0 * v : 0.0;
    else v = v * (1.0 / (double)(i - 1));
    for (int s = 0; s < a.nspecies; s++)
        if (a.qn[s]) v = v + a.qn[s][(size_t)k * ld + i];
    return v;
}

// Bz of mode array k from cu into b(3, k, .) (field_b_class.f03:760-796)
template <int K>
__device__ inline void bz_block(const FieldCommon &g, const double *cu, double *b, int k, double *stage,
                                double *red)
{
    const int nm = g.nm, ld = g.ld, nr = g.nr, t = threadIdx.x, m = mode_of(k);
    double v[K], d[1][K];
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads;
        v[kk] = (r < nr) ? bz_source(g, FieldArr{cu, nm, ld}, k, r + 1) : 0.0;
    }
    stage_in<K>(stage, v, nr, d[0]);
    const FactorT f[1] = {dfactor(g, FK_BZ, m)};
    seg_solve<K, 1>(f, d, red);
    stage_out<K>(stage, d[0], v);
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads;
        if (r < nr) FX(b, 2, k, r + 1) = (m > 0 && r == 0) ? 0.0 : v[kk];
    }
}

template <int K>
__global__ __launch_bounds__(kSolveThreads) void k_psi_bz(PsiBzArgs a)
{
    extern __shared__ double stage[];
    __shared__ double red[kRed];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int t = threadIdx.x;
    int b = blockIdx.x;
    if (b < nm) {
        if (!a.do_psi) return;
        const int k = b, m = mode_of(k);
        double v[K], d[1][K];
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const int r = t + kk * kSolveThreads, i = r + 1;
            double q = 0.0;
            if (r < nr) {
                if (a.from_partials) {
                    q = q_node(a, k, i);
                    a.q_spe[(size_t)k * ld + i] = q;
                } else {
                    q = a.q_spe[(size_t)k * ld + i];
                }
            }
            v[kk] = -1.0 * q;
        }
        if (t == 0 && a.from_partials) {
            a.q_spe[(size_t)k * ld] = q_node(a, k, 0);
            a.q_spe[(size_t)k * ld + nr + 1] = q_node(a, k, nr + 1);
        }
        stage_in<K>(stage, v, nr, d[0]);
        const FactorT f[1] = {dfactor(g, FK_PSI, m)};
        seg_solve<K, 1>(f, d, red);
        stage_out<K>(stage, d[0], v);
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const int r = t + kk * kSolveThreads;
            if (r < nr) a.psi[(size_t)k * ld + r + 1] = (m > 0 && r == 0) ? 0.0 : v[kk];
        }
        return;
    }
    b -= nm;
    if (b < nm) {
        if (!a.do_bz) return;
        bz_block<K>(g, a.cu, a.b_spe, b, stage, red);
        return;
    }
    // convergence_tester 'record' (simulation_class.f03:549-558)
    if (!a.do_record) return;
    if (t == 0) {
        a.it->done = 0;
        a.it->iters = 0;
        a.it->counter = 0;
    }
    for (int i = t + 1; i <= nr; i += blockDim.x) {
        double re = 0.0, im = 0.0;
        for (int mm = 0; mm <= g.M; mm++) {
            re = re + fabs(FX(a.b_spe, 1, kre(mm), i));
            if (mm > 0) im = im + fabs(FX(a.b_spe, 1, kim(mm), i));
        }
        a.conv_re[i] = re;
        a.conv_im[i] = im;
    }
}

static PsiBzArgs psibz_base(qpad_ctx *c)
{
    PsiBzArgs a;
    memset(&a, 0, sizeof(a));
    a.g = field_common(c);
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
        a.qn[s] = (s < c->nspecies && (c->sp[s].cfg.neutralized || c->sp[s].neutral)) ? c->sp[s].qn : nullptr;
        a.O0[s] = s < c->nspecies ? c->sp[s].O0 : nullptr;
        a.O1[s] = s < c->nspecies ? c->sp[s].O1 : nullptr;
        a.os[s] = s < c->nspecies ? c->sp[s].ov_start : nullptr;
    }
    return a;
}

static int run_psibz(qpad_ctx *c, const PsiBzArgs &a)
{
    int nblk = 2 * c->g.nm + 1;
    QPAD_DISPATCH_K(c->g.nr, k_psi_bz, dim3(nblk), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_psi_bz(qpad_ctx *c, bool do_psi, bool do_bz, bool do_record)
{
    PsiBzArgs a = psibz_base(c);
    a.do_psi = do_psi;
    a.do_bz = do_bz;
    a.do_record = do_record;
    a.from_partials = 1;
    return run_psibz(c, a);
}

int launch_psi_from_qspe(qpad_ctx *c)
{
    PsiBzArgs a = psibz_base(c);
    a.do_psi = 1;
    a.from_partials = 0;
    return run_psibz(c, a);
}

int launch_bz_only(qpad_ctx *c, int dst, int jay)
{
    PsiBzArgs a = psibz_base(c);
    a.cu = c->f[jay];
    a.b_spe = c->f[dst];
    a.do_bz = 1;
    return run_psibz(c, a);
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

// elementwise part at node i (0..nr+1) of mode array k
__device__ __forceinline__ void ezet_node(const EzEtArgs &a, int k, int i)
{
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
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
            double br = FX(a.bA, 0, k, i), bphi = FX(a.bA, 1, k, i);
            if (a.bB) {
                br = br + FX(a.bB, 0, k, i);
                bphi = bphi + FX(a.bB, 1, k, i);
            }
            et_node(g, br, bphi, FieldArr{a.psi, nm, ld}, k, i, ev);
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
            et_node(g, FX(a.bA, 0, k, i), FX(a.bA, 1, k, i), FieldArr{a.psi, nm, ld}, k, i, ev);
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

template <int K>
__global__ __launch_bounds__(kSolveThreads) void k_ez_et(EzEtArgs a)
{
    extern __shared__ double stage[];
    __shared__ double red[kRed];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int t = threadIdx.x;
    const int k = blockIdx.x, m = mode_of(k);
    if (a.check_done && a.it->done) return;
    const double *cu = a.jay;
    const double idr = g.idr, idrh = g.idrh;
    double v[K];
    double part = 0.0;
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads, i = r + 1;
        v[kk] = 0.0;
        if (r >= nr) continue;
        ezet_node(a, k, i);
        if (!cu) continue;
        // Ez source (field_e_class.f03:182-275)
        if (m == 0) {
            v[kk] = ez_source_0(g, FieldArr{cu, nm, ld}, i);
            if (i >= 2 && i <= nr - 2) part += v[kk] * (double)(i - 1);
        } else {
            v[kk] = ez_source_m(g, FieldArr{cu, nm, ld}, k, i);
        }
    }
    if (t == 0) {
        ezet_node(a, k, 0);
        ezet_node(a, k, nr + 1);
        if (a.final && k == 0) a.it->total_iters += a.it->iters;
    }
    if (!cu) return;
    if (m == 0) {
        double div = block_sum(part, red);
        // row 0 (axis) belongs to thread 0, kk = 0: discrete Gauss law
        if (t == 0) {
            div = div - ez_axis_edge(g, FieldArr{cu, nm, ld});
            v[0] = -8.0 * div;
        }
    }
    double d[1][K];
    stage_in<K>(stage, v, nr, d[0]);
    const FactorT f[1] = {dfactor(g, FK_EZ, m)};
    seg_solve<K, 1>(f, d, red);
    stage_out<K>(stage, d[0], v);
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads, i = r + 1;
        if (r < nr) {
            const double e = (m > 0 && i == 1) ? 0.0 : v[kk];
            FX(a.ez_dst, 2, k, i) = e;
            if (a.final) FX(a.h_e, 2, k, i) = e;
        }
    }
}

static EzEtArgs ezet_base(qpad_ctx *c)
{
    EzEtArgs a;
    memset(&a, 0, sizeof(a));
    a.g = field_common(c);
    a.psi = c->f[QPAD_F_PSI];
    a.it = c->it;
    return a;
}

static int run_ezet(qpad_ctx *c, const EzEtArgs &a)
{
    QPAD_DISPATCH_K(c->g.nr, k_ez_et, dim3(c->g.nm), 0, c->stream, a);
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
    const double *O0[kMaxSpecies], *O1[kMaxSpecies];
    const int32_t *os[kMaxSpecies];
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
        if (i >= 1 && i <= nr) val += cell_partial(a.P0[s], a.O0[s], a.os[s], i - 1, NV, v);
        if (i >= 2) val += cell_partial(a.P1[s], a.O1[s], a.os[s], i - 2, NV, v);
    }
    if (i == 0) {
        val = 0.0;
    } else if (i == 1) {
        // comp 0..2 = cu (r, phi, z), 3..4 = acu (r, phi), 5..7 = amu
        double fct;
        if (m == 0) fct = (comp == 2) ? 8.0 : 0.0;
        else if (m == 1) fct = (comp <= 1 || comp == 3 || comp == 4) ? 8.0 : 0.0;
        else if (m == 2) fct = (comp >= 5) ? 8.0 : 0.0;
        else fct = 0.0;
        val = (fct == 0.0) ? 0.0 : fct * val;
    } else {
        val = val * (1.0 / (double)(i - 1));
    }
    double *dst = (comp < 3) ? &FX(a.cu, comp, k, i)
                  : (comp < 5) ? &FX(a.acu, comp - 3, k, i) : &FX(a.amu, comp - 5, k, i);
    *dst = a.accumulate ? *dst + val : val;
}

static CombineJArgs combine_base(qpad_ctx *c)
{
    CombineJArgs a;
    memset(&a, 0, sizeof(a));
    a.nr = c->g.nr;
    a.ld = c->g.ld;
    a.nm = c->g.nm;
    a.cu = c->f[QPAD_F_CU];
    a.acu = c->f[QPAD_F_ACU];
    a.amu = c->f[QPAD_F_AMU];
    a.it = c->it;
    return a;
}

static int run_combine(qpad_ctx *c, const CombineJArgs &a)
{
    int64_t total = (int64_t)8 * c->g.nm * (c->g.nr + 2);
    hipLaunchKernelGGL(k_combine_j, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_combine_j(qpad_ctx *c, bool check_done)
{
    CombineJArgs a = combine_base(c);
    a.nspecies = c->nspecies;
    for (int s = 0; s < kMaxSpecies; s++) {
        a.P0[s] = s < c->nspecies ? c->sp[s].P0 : nullptr;
        a.P1[s] = s < c->nspecies ? c->sp[s].P1 : nullptr;
        a.O0[s] = s < c->nspecies ? c->sp[s].O0 : nullptr;
        a.O1[s] = s < c->nspecies ? c->sp[s].O1 : nullptr;
        a.os[s] = s < c->nspecies ? c->sp[s].ov_start : nullptr;
    }
    a.check_done = check_done;
    return run_combine(c, a);
}

int launch_combine_j_species(qpad_ctx *c, int sp)
{
    CombineJArgs a = combine_base(c);
    a.nspecies = 1;
    a.P0[0] = c->sp[sp].P0;
    a.P1[0] = c->sp[sp].P1;
    a.O0[0] = c->sp[sp].O0;
    a.O1[0] = c->sp[sp].O1;
    a.os[0] = c->sp[sp].ov_start;
    a.accumulate = 1;
    return run_combine(c, a);
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

template <int K>
__global__ __launch_bounds__(kSolveThreads) void k_bperp_bz(BperpArgs a)
{
    extern __shared__ double stage[];
    __shared__ double red[kRed];
    __shared__ int s_last;
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int t = threadIdx.x;
    if (a.check_done && a.it->done) return;
    const int b = blockIdx.x;
    if (b <= M) {
        if (a.compare_only) {
            conv_compare_last<K>(g, a.b_spe, a.conv_re, a.conv_im, a.it, a.reltol, a.abstol, a.iter_max, red,
                                 &s_last);
            return;
        }
        if (!a.do_bperp && !a.do_djdxi_only) return;
        const int m = b, kr = kre(m), ki = kim(m);
        double src[4][K];
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const int r = t + kk * kSolveThreads, i = r + 1;
            src[0][kk] = 0.0; src[1][kk] = 0.0; src[2][kk] = 0.0; src[3][kk] = 0.0;
            if (r >= nr) continue;
            double dr_[2], di_[2] = {0.0, 0.0};
            const FieldArr acu{a.acu, nm, ld}, amu{a.amu, nm, ld};
            djdxi_node(g, acu, amu, kr, i, dr_);
            FX(a.dcu, 0, kr, i) = dr_[0];
            FX(a.dcu, 1, kr, i) = dr_[1];
            if (m > 0) {
                djdxi_node(g, acu, amu, ki, i, di_);
                FX(a.dcu, 0, ki, i) = di_[0];
                FX(a.dcu, 1, ki, i) = di_[1];
            }
            if (a.do_djdxi_only) continue;
            double o[4];
            bpm_source(g, FieldArr{a.cu, nm, ld}, FieldArr{a.b_spe, nm, ld}, m, i, dr_, di_, o);
            src[0][kk] = o[0]; src[1][kk] = o[1]; src[2][kk] = o[2]; src[3][kk] = o[3];
        }
        if (a.do_djdxi_only) return;
        double d[4][K];
#pragma unroll
        for (int s = 0; s < 4; s++) stage_in<K>(stage, src[s], nr, d[s]);
        const FactorT fp = dfactor(g, FK_BPLUS, m), fm = dfactor(g, FK_BMINUS, m);
        const FactorT f[4] = {fp, fp, fm, fm};
        seg_solve<K, 4>(f, d, red);
#pragma unroll
        for (int s = 0; s < 4; s++) stage_out<K>(stage, d[s], src[s]);
        // recombination (field_b_class.f03:715-752); every b_spe(1:2, m) read
        // above happened before the solve's barriers
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const int r = t + kk * kSolveThreads, i = r + 1;
            if (r >= nr) continue;
            const double sv[4] = {src[0][kk], src[1][kk], src[2][kk], src[3][kk]};
            double o[4];
            bpm_recombine(m, i, sv, o);
            FX(a.b_spe, 0, kr, i) = o[0];
            FX(a.b_spe, 1, kr, i) = o[2];
            if (m > 0) {
                FX(a.b_spe, 0, ki, i) = o[1];
                FX(a.b_spe, 1, ki, i) = o[3];
            }
        }
        if (a.do_compare)
            conv_compare_last<K>(g, a.b_spe, a.conv_re, a.conv_im, a.it, a.reltol, a.abstol, a.iter_max, red,
                                 &s_last);
        return;
    }
    const int k = b - (M + 1);
    if (k < nm && a.do_bz) bz_block<K>(g, a.cu, a.b_spe, k, stage, red);
}

static BperpArgs bperp_base(qpad_ctx *c)
{
    BperpArgs a;
    memset(&a, 0, sizeof(a));
    a.g = field_common(c);
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

static int run_bperp(qpad_ctx *c, const BperpArgs &a)
{
    int nblk = (c->g.M + 1) + c->g.nm;
    QPAD_DISPATCH_K(c->g.nr, k_bperp_bz, dim3(nblk), 0, c->stream, a);
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

// ---------------------------------------------------------------------------
// Beam B_perp (field_b_class.f03:309-358, :545-701), batched over slices
// ---------------------------------------------------------------------------
struct BtBeamArgs {
    FieldCommon g;
    const double *hq;   // first q_beam slice to solve ([1][nm][ld] per slice)
    double *hb;         // first b_beam slice to write ([3][nm][ld] per slice)
    size_t qstride, bstride;
};

template <int K>
__global__ __launch_bounds__(kSolveThreads) void k_bt_beam(BtBeamArgs a)
{
    extern __shared__ double stage[];
    __shared__ double red[kRed];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int t = threadIdx.x;
    const int m = blockIdx.y;
    const int kr = kre(m), ki = kim(m);
    const double *q = a.hq + (size_t)blockIdx.x * a.qstride;
    double *bf = a.hb + (size_t)blockIdx.x * a.bstride;
    double v[2][K], d[2][K];
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads, i = r + 1;
        v[0][kk] = (r < nr) ? -1.0 * q[(size_t)kr * ld + i] : 0.0;
        v[1][kk] = (r < nr && m > 0) ? -1.0 * q[(size_t)ki * ld + i] : 0.0;
    }
    stage_in<K>(stage, v[0], nr, d[0]);
    stage_in<K>(stage, v[1], nr, d[1]);
    const FactorT fb = dfactor(g, FK_BT, m);
    const FactorT f[2] = {fb, fb};
    seg_solve<K, 2>(f, d, red);
    // post-processing needs neighbouring rows: buf(i) = stage[pidx(i - 1)]
    const double idr = g.idr, idrh = g.idrh, dm = (double)m;
#define BUF(i) stage[pidx<K>((i) - 1)]
    // real part of the potential: B_phi(Re), B_r(Im)
    blocked_to_stage<K>(stage, d[0]);
    __syncthreads();
    for (int i = t + 1; i <= nr; i += kSolveThreads) {
        double r1v, i0v;
        if (i == 1) {
            r1v = (m == 1) ? -idr * BUF(2) : 0.0;
            i0v = (m == 1) ? idr * dm * BUF(2) : 0.0;
        } else if (i == nr) {
            r1v = -idrh * (3.0 * BUF(nr) - 4.0 * BUF(nr - 1) + BUF(nr - 2));
            i0v = idr / (double)(nr - 1) * dm * BUF(nr);
        } else {
            r1v = -idrh * (BUF(i + 1) - BUF(i - 1));
            i0v = idr / (double)(i - 1) * dm * BUF(i);
        }
        if (m == 0) {
            FX(bf, 0, 0, i) = 0.0;
            FX(bf, 1, 0, i) = r1v;
        } else {
            FX(bf, 1, kr, i) = r1v;
            FX(bf, 0, ki, i) = i0v;
        }
    }
    if (m == 0) return;
    __syncthreads();
    blocked_to_stage<K>(stage, d[1]);
    __syncthreads();
    for (int i = t + 1; i <= nr; i += kSolveThreads) {
        double r0v, i1v;
        if (i == 1) {
            r0v = (m == 1) ? -idr * dm * BUF(2) : 0.0;
            i1v = (m == 1) ? -idr * BUF(2) : 0.0;
        } else if (i == nr) {
            r0v = -(idr / (double)(nr - 1)) * dm * BUF(nr);
            i1v = -idrh * (3.0 * BUF(nr) - 4.0 * BUF(nr - 1) + BUF(nr - 2));
        } else {
            r0v = -(idr / (double)(i - 1)) * dm * BUF(i);
            i1v = -idrh * (BUF(i + 1) - BUF(i - 1));
        }
        FX(bf, 0, kr, i) = r0v;
        FX(bf, 1, ki, i) = i1v;
    }
#undef BUF
}

int launch_bt_beam_all(qpad_ctx *c)
{
    BtBeamArgs a;
    a.g = field_common(c);
    a.qstride = (size_t)c->g.nm * c->g.ld;
    a.bstride = (size_t)3 * c->g.nm * c->g.ld;
    a.hq = c->h_q_beam + a.qstride;   // slices 1..nxi
    a.hb = c->h_b_beam + a.bstride;
    QPAD_DISPATCH_K(c->g.nr, k_bt_beam, dim3(c->g.nxi, c->g.M + 1), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_bt_beam_one(qpad_ctx *c, int j)
{
    BtBeamArgs a;
    a.g = field_common(c);
    a.qstride = 0;
    a.bstride = 0;
    a.hq = c->h_q_beam + (size_t)j * c->g.nm * c->g.ld;
    a.hb = c->f[QPAD_F_B_BEAM];
    // q_beam%copy_slice(j, p_copy_2to1)
    QPAD_HIP_CHECK(hipMemcpyAsync(c->f[QPAD_F_Q_BEAM], a.hq, sizeof(double) * c->g.nm * c->g.ld,
                                  hipMemcpyDeviceToDevice, c->stream));
    QPAD_DISPATCH_K(c->g.nr, k_bt_beam, dim3(1, c->g.M + 1), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// the LDS stage of the K = 16 variants exceeds the default dynamic LDS
template <int K>
static int set_lds_attr()
{
    const int bytes = 8 * stage_doubles<K>();
    const void *ks[] = {(const void *)k_psi_bz<K>, (const void *)k_ez_et<K>, (const void *)k_bperp_bz<K>,
                        (const void *)k_bt_beam<K>};
    for (const void *k : ks)
        QPAD_HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return 0;
}

int field_kernels_init(size_t)
{
    QPAD_TRY(set_lds_attr<1>());
    QPAD_TRY(set_lds_attr<2>());
    QPAD_TRY(set_lds_attr<4>());
    QPAD_TRY(set_lds_attr<8>());
    QPAD_TRY(set_lds_attr<16>());
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
