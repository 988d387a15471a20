// k_slice.hip -- the fused B_perp predictor-corrector iteration of a slice
// (simulation_class.f03:370-398) after the current deposit:
//
//   k_src    every node and mode in parallel: cu, acu, amu from the per-cell
//            deposit partials (written out), dJ/dxi (written out), and the
//            right-hand sides of every solve of the iteration (B+/B-, Bz,
//            Ez) into the `src` buffer, rows interleaved for coalescing;
//   k_solve  one workgroup per system: B_perp (4 systems per mode, then the
//            recombination, E_perp and b = b_spe + b_beam of that mode), Bz,
//            Ez; the last B_perp workgroup runs the convergence test.
//
// E and b written by k_solve are the fields of the next iteration's deposit
// (the reference recomputes them at the top of that iteration from the same
// cu, b_spe and psi).  k_epilogue is the slice's closing copy-out
// (simulation_class.f03:401-456).  Every kernel returns at once when the
// device-side `done` flag says the iteration has converged, so iterations
// are enqueued speculatively with no host round trip.
#include <algorithm>
#include <cstring>

#include "field_src.h"
#include "kernels.h"
#include "conv.h"

namespace qpad {

// threads per workgroup of the once-per-slice node kernels (k_start_src,
// k_eperp, k_epilogue: a thread per node and mode array); small workgroups
// spread the few thousand nodes over many CUs
constexpr int kNodeThreads = 256;
inline int node_blocks(int nr) { return (nr + 2 + kNodeThreads - 1) / kNodeThreads; }

// src buffer rows: B+/B- of mode m at 4m + s (s = Re B+, Im B+, Re B-,
// Im B-), Bz of mode array k at 4(M+1) + k, Ez at 4(M+1) + nm + k, psi at
// 4(M+1) + 2nm + k; row r = i - 1 of each system at [sys * nr + r]
__host__ __device__ inline int src_bz(int M, int k) { return 4 * (M + 1) + k; }
__host__ __device__ inline int src_ez(int M, int nm, int k) { return 4 * (M + 1) + nm + k; }
__host__ __device__ inline int src_psi(int M, int nm, int k) { return 4 * (M + 1) + 2 * nm + k; }
__host__ __device__ inline int src_rows(int M, int nm) { return 4 * (M + 1) + 3 * nm; }

struct SrcArgs {
    FieldCommon g;
    const double *P0[kMaxSpecies], *P1[kMaxSpecies];
    // overflow rows of crowded cells (cell_partial)
    const double *O0[kMaxSpecies], *O1[kMaxSpecies];
    const int32_t *os[kMaxSpecies];
    int ns;
    double *cu_out, *acu_out, *amu_out, *dcu;
    const double *b_spe;
    double *src;
    const IterState *it;
    int late_done;   // 1: test the converged flag after the first loads
};

// nodes per k_src workgroup; its LDS tile also holds a halo of 2 nodes on
// each side (the sources use nodes i-2 .. i+1)
constexpr int kSrcTile = 16;
constexpr int kSrcHalo = 2;
constexpr int kSrcTL = kSrcTile + 2 * kSrcHalo;
static_assert(kSrcTile * (kMaxMode + 1) <= 256, "one B+/B- item per k_src thread");

// cu (base 0), acu (3) or amu (5) of the tile in LDS as [comp8][k][j],
// j = i - (i0 - 2)
struct TileArr {
    const double *s;
    int nm, base, i0;
    __device__ __forceinline__ double operator()(int c, int k, int i) const
    {
        return s[((size_t)(base + c) * nm + k) * kSrcTL + (i - i0 + kSrcHalo)];
    }
};

// cu, acu, amu at node i (0 <= i <= nr+1) from the per-cell partials of
// every species, with the axis rules and 1/r of amjdeposit_robust_part2d
// (part2d_class.f03:914-1004).  P0[cell][v] is the deposit of cell `cell`
// on its inner node, P1 on its outer node, v = comp8 * nm + k.  Split in an
// address step (clamped, so the loads are always valid) and a combine step.
__device__ __forceinline__ void partial_addr(const FieldCommon &g, int comp, int k, int i, int &cell0,
                                             int &cell1, size_t &c0, size_t &c1)
{
    const int nm = g.nm, nr = g.nr, NV = 8 * nm, v = comp * nm + k;
    cell0 = min(max(i - 1, 0), nr - 1);
    cell1 = min(max(i - 2, 0), nr - 1);
    c0 = (size_t)cell0 * NV + v;
    c1 = (size_t)cell1 * NV + v;
}

template <int NS>
__device__ __forceinline__ double combine_node(int nr, int comp, int k, int i, const double (&d0)[NS],
                                               const double (&d1)[NS])
{
    const int m = (k + 1) >> 1;
    const bool in0 = i >= 1 && i <= nr, in1 = i >= 2;
    double val = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (in0) val += d0[s];
        if (in1) val += d1[s];
    }
    double fct;
    if (m == 0) fct = (comp == 2) ? 8.0 : 0.0;
    else if (m == 1) fct = (comp <= 1 || comp == 3 || comp == 4) ? 8.0 : 0.0;
    else if (m == 2) fct = (comp >= 5) ? 8.0 : 0.0;
    else fct = 0.0;
    const double axis = (fct == 0.0) ? 0.0 : fct * val;
    const double inner = val * (1.0 / (double)max(i - 1, 1));
    return i == 0 ? 0.0 : (i == 1 ? axis : inner);
}

// a uniform int read with a vector load (the zero offset comes from an asm,
// so the compiler cannot turn it into a scalar load)
__device__ __forceinline__ int flag_vload(const int *p)
{
    int z;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(z));
    return p[z];
}

// b_spe components 0, 1 of mode arrays kr and ki (= kr for m = 0) at one
// node, as bpm_source reads them (f3(c, k, i) at that node only)
struct BspeNode {
    int kr;
    double r0, r1, i0, i1;
    __device__ __forceinline__ double operator()(int c, int k, int) const
    {
        return k == kr ? (c == 0 ? r0 : r1) : (c == 0 ? i0 : i1);
    }
};

// one workgroup per tile of kSrcTile nodes: (1) cu, acu, amu of the tile
// and its halo into LDS; (2) B+/B- sources of every mode and dJ/dxi at the
// tile's nodes 1..nr; (3) cu, acu, amu out (nodes 0..nr+1) and the Bz and Ez
// sources of every mode array
template <int NS>
__global__ __launch_bounds__(256) void k_src(SrcArgs a)
{
    extern __shared__ double tile[];
    // the converged flag: with adaptive queuing (late_done) almost every
    // launch works, and the flag is tested after the first pass's loads are
    // issued (one round trip less); with fixed queuing about half the
    // launches come after convergence and return before loading anything
    // (a vector load, through an offset the compiler cannot see is zero: its
    // wait does not also wait for the kernel-argument scalar loads, and it
    // retires before the data loads issued after it)
    int done = flag_vload(&a.it->done);   // requested first in both modes
    if (!a.late_done) {
        __asm__ volatile("");   // keeps this test a branch of its own
        if (__builtin_amdgcn_readfirstlane(done)) return;
    }
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int t = threadIdx.x, T = blockDim.x;
    const int i0 = blockIdx.x * kSrcTile;
    const int n_tile = 8 * nm * kSrcTL;
    // does any cell the tile reads have overflow rows (crowded cells)?  The
    // test is uniform over the workgroup and is only needed after the
    // pass's loads are issued.  Such tiles take a lean path (one entry at a
    // time, each cell's overflow rows added in row order), so that the
    // common path keeps its registers and occupancy.
    const int lo = max(i0 - kSrcHalo - 2, 0), hi = min(i0 + kSrcTile + kSrcHalo - 2, nr - 1);
    int os_lo[NS], os_hi[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
        os_lo[q] = a.os[q][lo];
        os_hi[q] = a.os[q][hi + 1];
    }
    // b_spe at the node of this thread's B+/B- item (the only b_spe values
    // the sources read), requested first; one item per thread
    // (kSrcTile * (kMaxMode + 1) <= 256 threads)
    const int m2 = t / kSrcTile, i2 = i0 + t % kSrcTile;
    const bool item2 = t < kSrcTile * (M + 1) && i2 >= 1 && i2 <= nr;
    const int kr2 = m2 == 0 ? 0 : 2 * m2 - 1, ki2 = m2 == 0 ? 0 : 2 * m2;
    const int i2c = min(max(i2, 1), nr);
    const BspeNode bn{kr2, FX(a.b_spe, 0, min(kr2, nm - 1), i2c), FX(a.b_spe, 1, min(kr2, nm - 1), i2c),
                      FX(a.b_spe, 0, min(ki2, nm - 1), i2c), FX(a.b_spe, 1, min(ki2, nm - 1), i2c)};
    const int NV = 8 * nm;
    // 8 entries per thread per pass so that their loads are in flight
    // together; entry idx = j * NV + v (node j of the tile, value v), so that
    // neighbouring threads read neighbouring values of one cell's partial row
    // (coalesced) and the tile in LDS is [v][j]
    double d0[8][NS], d1[8][NS];
    auto pass_loads = [&](int base) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = min(base + u * T, n_tile - 1);
            const int j = idx / NV, vv = idx - j * NV, comp = vv / nm, k = vv - comp * nm;
            const int i = min(max(i0 - kSrcHalo + j, 0), nr + 1);
            int cl0, cl1;
            size_t c0, c1;
            partial_addr(g, comp, k, i, cl0, cl1, c0, c1);
#pragma unroll
            for (int q = 0; q < NS; q++) {
                d0[u][q] = a.P0[q][c0];
                d1[u][q] = a.P1[q][c1];
            }
        }
    };
    auto pass_combine = [&](int base, bool ovf) {
        if (!ovf) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * T;
                if (idx >= n_tile) break;
                const int j = idx / NV, vv = idx - j * NV, comp = vv / nm, k = vv - comp * nm;
                const int i = i0 - kSrcHalo + j;
                const double v = combine_node<NS>(nr, comp, k, min(max(i, 0), nr + 1), d0[u], d1[u]);
                tile[vv * kSrcTL + j] = (i >= 0 && i <= nr + 1) ? v : 0.0;
            }
        } else {
#pragma unroll 1
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * T;
                if (idx >= n_tile) break;
                const int j = idx / NV, vv = idx - j * NV, comp = vv / nm, k = vv - comp * nm;
                const int i = i0 - kSrcHalo + j;
                const int ic = min(max(i, 0), nr + 1);
                int cl0, cl1;
                size_t c0, c1;
                partial_addr(g, comp, k, ic, cl0, cl1, c0, c1);
                double e0[NS], e1[NS];
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    e0[q] = add_rows(a.P0[q][c0], a.O0[q], a.os[q][cl0], a.os[q][cl0 + 1], NV, vv);
                    e1[q] = add_rows(a.P1[q][c1], a.O1[q], a.os[q][cl1], a.os[q][cl1 + 1], NV, vv);
                }
                const double val = combine_node<NS>(nr, comp, k, ic, e0, e1);
                tile[vv * kSrcTL + j] = (i >= 0 && i <= nr + 1) ? val : 0.0;
            }
        }
    };
    // first pass peeled (a loop header would wait for every load in flight,
    // the b_spe requests above included, before this pass's loads)
    pass_loads(t);
    // (the empty asms keep the compiler from hoisting the tests, and the
    // wait for their loads, above this pass's loads)
    if (a.late_done) {
        __asm__ volatile("" : "+v"(done));
        if (__builtin_amdgcn_readfirstlane(done)) return;
    }
    bool ovf = false;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        int l_ = os_lo[q], h_ = os_hi[q];
        __asm__ volatile("" : "+v"(l_), "+v"(h_));
        ovf = ovf || __builtin_amdgcn_readfirstlane(h_) != __builtin_amdgcn_readfirstlane(l_);
    }
    pass_combine(t, ovf);
    for (int base = t + 8 * T; base < n_tile; base += 8 * T) {
        pass_loads(base);
        pass_combine(base, ovf);
    }
    __syncthreads();
    const TileArr cu{tile, nm, 0, i0}, acu{tile, nm, 3, i0}, amu{tile, nm, 5, i0};
    // B+/B- of mode m and dJ/dxi
    if (item2) {
        const int m = m2, i = i2;
        const int kr = kr2, ki = 2 * m;
        double dr_[2], di_[2] = {0.0, 0.0}, bpm[4];
        djdxi_node(g, acu, amu, kr, i, dr_);
        if (m > 0) djdxi_node(g, acu, amu, ki, i, di_);
        bpm_source(g, cu, bn, m, i, dr_, di_, bpm);
        FX(a.dcu, 0, kr, i) = dr_[0];
        FX(a.dcu, 1, kr, i) = dr_[1];
        if (m > 0) {
            FX(a.dcu, 0, ki, i) = di_[0];
            FX(a.dcu, 1, ki, i) = di_[1];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) a.src[(size_t)(4 * m + q) * nr + i - 1] = bpm[q];
    }
    // cu, acu, amu out; Bz and Ez sources
    for (int it = t; it < kSrcTile * nm; it += T) {
        const int k = it / kSrcTile, i = i0 + it % kSrcTile, m = mode_of(k);
        if (i > nr + 1) continue;
#pragma unroll
        for (int c = 0; c < 3; c++) FX(a.cu_out, c, k, i) = cu(c, k, i);
#pragma unroll
        for (int c = 0; c < 2; c++) FX(a.acu_out, c, k, i) = acu(c, k, i);
#pragma unroll
        for (int c = 0; c < 3; c++) FX(a.amu_out, c, k, i) = amu(c, k, i);
        if (i < 1 || i > nr) continue;
        a.src[(size_t)src_bz(M, k) * nr + i - 1] = bz_source(g, cu, k, i);
        a.src[(size_t)src_ez(M, nm, k) * nr + i - 1] =
            (m == 0) ? ez_source_0(g, cu, i) : ez_source_m(g, cu, k, i);
    }
}

struct SolveArgs {
    FieldCommon g;
    double *src;             // B+/B- solutions are written back into their rows
    double *b_spe, *e;
    const double *cu;
    const IterState *it;
    int late_done;   // as in SrcArgs
};

// rows r = t + kk * kSolveThreads of a system (clamped loads, so that all K
// are in flight together; rows past nr read as 0)
template <int K>
__device__ inline void load_rows(const double *p, int nr, double (&v)[K])
{
    double x[K];
#pragma unroll
    for (int kk = 0; kk < K; kk++) x[kk] = p[min((int)threadIdx.x + kk * kSolveThreads, nr - 1)];
#pragma unroll
    for (int kk = 0; kk < K; kk++) v[kk] = ((int)threadIdx.x + kk * kSolveThreads < nr) ? x[kk] : 0.0;
}

// one workgroup per system: 4 (M+1) B+/B- systems (field_b_class.f03:
// 388-497; the m = 0 imaginary ones are empty and skipped), then Bz and Ez
// of every mode array
template <int K>
__global__ __launch_bounds__(kSolveThreads) void k_solve(SolveArgs a)
{
    extern __shared__ double stage[];
    __shared__ double red[kRed];
    // tested once the rows are requested (late_done); a vector load
    int done = flag_vload(&a.it->done);
    if (!a.late_done) {
        __asm__ volatile("");
        if (__builtin_amdgcn_readfirstlane(done)) return;
    }
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    int row, kind, m, role;
    if (b < 4 * (M + 1)) {
        m = b >> 2;
        if (m == 0 && (b & 1)) return;
        row = b;
        kind = (b & 2) ? FK_BMINUS : FK_BPLUS;
        role = 0;
    } else if (b < 4 * (M + 1) + nm) {
        const int k = b - 4 * (M + 1);
        m = mode_of(k);
        row = src_bz(M, k);
        kind = FK_BZ;
        role = 1;
    } else {
        const int k = b - 4 * (M + 1) - nm;
        m = mode_of(k);
        row = src_ez(M, nm, k);
        kind = FK_EZ;
        role = 2;
    }
    double v[K], d[1][K];
    load_rows<K>(a.src + (size_t)row * nr, nr, v);
    // the factor rows are requested together with the right-hand side
    const FactorT f[1] = {dfactor(g, kind, m)};
    FactorRows<K, 1> fr;
    load_factors<K, 1>(f, fr);
    if (a.late_done) {
        __asm__ volatile("" : "+v"(done));
        if (__builtin_amdgcn_readfirstlane(done)) return;
    }
    if (role == 2 && m == 0) {
        // discrete Gauss law on the axis row of Ez
        double part = 0.0;
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const int i = t + kk * kSolveThreads + 1;
            if (i >= 2 && i <= nr - 2) part += v[kk] * (double)(i - 1);
        }
        double div = block_sum(part, red);
        if (t == 0) {
            div = div - ez_axis_edge(g, FieldArr{a.cu, nm, ld});
            v[0] = -8.0 * div;
        }
    }
    stage_in<K>(stage, v, nr, d[0]);
    seg_solve_rows<K, 1>(fr, d, red);
    stage_out<K>(stage, d[0], v);
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads, i = r + 1;
        if (r >= nr) continue;
        if (role == 0) {
            a.src[(size_t)row * nr + r] = v[kk];
        } else {
            const int k = role == 1 ? b - 4 * (M + 1) : b - 4 * (M + 1) - nm;
            const double x = (m > 0 && i == 1) ? 0.0 : v[kk];
            if (role == 1) FX(a.b_spe, 2, k, i) = x;
            else FX(a.e, 2, k, i) = x;
        }
    }
}

struct BfieldArgs {
    FieldCommon g;
    const double *src, *b_beam, *psi;
    double *b_spe, *b_out, *e;
    double *conv_re, *conv_im;
    double *conv_blk;   // this iteration's half: [workgroup][max old^2, max diff^2]
    IterState *it;
    int l;              // iteration index (1-based)
    int late_done;      // as in SrcArgs
};

// thread per node i = 0..nr+1, all modes: B_r, B_phi of b_spe from the
// B+/B- solutions (field_b_class.f03:715-752), b = b_spe + b_beam, E_perp
// (field_e_class.f03:443-504), and convergence_tester 'compare'
// (simulation_class.f03:560-599): the node's sum_m |B_phi,m| against the
// recorded one, max over the workgroup's nodes (the next iteration's deposit
// reduces the workgroup maxima).  Loads are issued for all modes before any
// arithmetic.
static_assert((16 * kSolveThreads + 2 + kBfThreads - 1) / kBfThreads <= kConvMaxBlocks,
              "the next deposit reduces every B_perp workgroup's maxima");

template <int M>
__global__ __launch_bounds__(kBfThreads) void k_bfields(BfieldArgs a)
{
    constexpr int NM = 2 * M + 1;
    __shared__ double red[8];
    // tested once the loads are requested (late_done); a vector load
    int done = flag_vload(&a.it->done);
    if (!a.late_done) {
        __asm__ volatile("");
        if (__builtin_amdgcn_readfirstlane(done)) return;
    }
    const FieldCommon &g = a.g;
    const int nm = NM, ld = g.ld, nr = g.nr;
    const int t = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + t;
    const bool valid = i <= nr + 1;
    const int ic = min(i, nr + 1);
    const bool inner = valid && ic >= 1 && ic <= nr;
    const int r = min(max(ic - 1, 0), nr - 1);
    // ---- loads ----
    double raw[M + 1][4], bb[NM][3], bs[NM][3], pw[NM][4], pp_i[NM], pp_2[NM];
#pragma unroll
    for (int m = 0; m <= M; m++)
#pragma unroll
        for (int q = 0; q < 4; q++) raw[m][q] = a.src[(size_t)(4 * m + q) * nr + r];
#pragma unroll
    for (int k = 0; k < NM; k++) {
        const int p = k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bb[k][c] = FX(a.b_beam, c, k, ic);
            bs[k][c] = FX(a.b_spe, c, k, ic);
        }
#pragma unroll
        for (int w = 0; w < 4; w++) pw[k][w] = a.psi[(size_t)k * ld + min(max(ic - 2 + w, 0), nr + 1)];
        pp_i[k] = a.psi[(size_t)p * ld + ic];
        pp_2[k] = a.psi[(size_t)p * ld + 2];
    }
    const double ore = a.conv_re[ic], oim = a.conv_im[ic];
    if (a.late_done) {
        __asm__ volatile("" : "+v"(done));
        if (__builtin_amdgcn_readfirstlane(done)) return;
    }
    // ---- b_spe (inner nodes), b, E_perp ----
#pragma unroll
    for (int m = 0; m <= M; m++) {
        if (!inner) break;
        const int kr = m == 0 ? 0 : 2 * m - 1, ki = 2 * m;
        const double sv[4] = {raw[m][0], m ? raw[m][1] : 0.0, raw[m][2], m ? raw[m][3] : 0.0};
        double o[4];
        bpm_recombine(m, ic, sv, o);
        bs[kr][0] = o[0];
        bs[kr][1] = o[2];
        if (m > 0) {
            bs[ki][0] = o[1];
            bs[ki][1] = o[3];
        }
    }
    double sre = 0.0, sim = 0.0, dre = ore, dim = oim;
#pragma unroll
    for (int m = 0; m <= M; m++) {
        const double vr = fabs(bs[m == 0 ? 0 : 2 * m - 1][1]);
        dre = dre - vr;
        sre = sre + vr;
        if (m > 0) {
            const double vi = fabs(bs[2 * m][1]);
            dim = dim - vi;
            sim = sim + vi;
        }
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < NM; k++) {
            const double br = bs[k][0] + bb[k][0], bphi = bs[k][1] + bb[k][1];
            FX(a.b_out, 0, k, ic) = br;
            FX(a.b_out, 1, k, ic) = bphi;
            FX(a.b_out, 2, k, ic) = bs[k][2] + bb[k][2];
            if (!inner) continue;
            FX(a.b_spe, 0, k, ic) = bs[k][0];
            FX(a.b_spe, 1, k, ic) = bs[k][1];
            double ev[2];
            et_window(g, br, bphi, pw[k][0], pw[k][1], pw[k][2], pw[k][3], pp_i[k], pp_2[k], k, ic, ev);
            FX(a.e, 0, k, ic) = ev[0];
            FX(a.e, 1, k, ic) = ev[1];
        }
    }
    double mx_old = -INFINITY, mx_d = -INFINITY;
    if (inner) {
        mx_old = ore * ore + oim * oim;
        mx_d = dre * dre + dim * dim;
        a.conv_re[ic] = sre;
        a.conv_im[ic] = sim;
    }
    mx_old = block_max(mx_old, red);
    mx_d = block_max(mx_d, red);
    if (t == 0) {
        a.conv_blk[2 * blockIdx.x] = mx_old;
        a.conv_blk[2 * blockIdx.x + 1] = mx_d;
        if (blockIdx.x == 0) a.it->iters = a.l;
    }
}

int launch_iteration_fields(qpad_ctx *c, int j, int l)
{
    const FieldCommon g = field_common(c);
    SrcArgs s;
    memset(&s, 0, sizeof(s));
    s.g = g;
    s.ns = c->nspecies;
    for (int q = 0; q < c->nspecies; q++) {
        s.P0[q] = c->sp[q].P0;
        s.P1[q] = c->sp[q].P1;
        s.O0[q] = c->sp[q].O0;
        s.O1[q] = c->sp[q].O1;
        s.os[q] = c->sp[q].ov_start;
    }
    s.cu_out = c->f[QPAD_F_CU];
    s.acu_out = c->f[QPAD_F_ACU];
    s.amu_out = c->f[QPAD_F_AMU];
    s.dcu = c->f[QPAD_F_DCU];
    s.b_spe = c->f[QPAD_F_B_SPE];
    s.src = c->src;
    s.it = c->it;
    s.late_done = c->adaptive;
    const size_t src_lds = sizeof(double) * 8 * c->g.nm * kSrcTL;
    const dim3 sgrid((unsigned)((c->g.nr + 2 + kSrcTile - 1) / kSrcTile));
    switch (c->nspecies) {
    case 1: hipLaunchKernelGGL(k_src<1>, sgrid, dim3(256), src_lds, c->stream, s); break;
    case 2: hipLaunchKernelGGL(k_src<2>, sgrid, dim3(256), src_lds, c->stream, s); break;
    case 3: hipLaunchKernelGGL(k_src<3>, sgrid, dim3(256), src_lds, c->stream, s); break;
    default: hipLaunchKernelGGL(k_src<4>, sgrid, dim3(256), src_lds, c->stream, s); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());

    SolveArgs a;
    a.g = g;
    a.src = c->src;
    a.b_spe = c->f[QPAD_F_B_SPE];
    a.e = c->f[QPAD_F_E];
    a.cu = c->f[QPAD_F_CU];
    a.it = c->it;
    a.late_done = c->adaptive;
    const int nblk = 4 * (c->g.M + 1) + 2 * c->g.nm;
    QPAD_DISPATCH_K(c->g.nr, k_solve, dim3(nblk), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());

    BfieldArgs f;
    f.g = g;
    f.src = c->src;
    f.b_beam = c->h_b_beam + (size_t)j * 3 * c->g.nm * c->g.ld;
    f.psi = c->f[QPAD_F_PSI];
    f.b_spe = c->f[QPAD_F_B_SPE];
    f.b_out = c->f[QPAD_F_B];
    f.e = c->f[QPAD_F_E];
    f.conv_re = c->conv_re;
    f.conv_im = c->conv_im;
    f.conv_blk = c->conv_blk + (size_t)(l & 1) * 2 * conv_blocks(c->g.nr);
    f.it = c->it;
    f.l = l;
    f.late_done = c->adaptive;
    const dim3 fgrid((unsigned)conv_blocks(c->g.nr)), fblk(kBfThreads);
    switch (c->g.M) {
    case 0: hipLaunchKernelGGL(k_bfields<0>, fgrid, fblk, 0, c->stream, f); break;
    case 1: hipLaunchKernelGGL(k_bfields<1>, fgrid, fblk, 0, c->stream, f); break;
    case 2: hipLaunchKernelGGL(k_bfields<2>, fgrid, fblk, 0, c->stream, f); break;
    case 3: hipLaunchKernelGGL(k_bfields<3>, fgrid, fblk, 0, c->stream, f); break;
    default: hipLaunchKernelGGL(k_bfields<4>, fgrid, fblk, 0, c->stream, f); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// slice epilogue (simulation_class.f03:401-456): cbq, E of the plasma alone,
// history copies, beam fields of slice j into the slice arrays, J_perp
// predictor
// ---------------------------------------------------------------------------
struct EpiArgs {
    FieldCommon g;
    const double *e, *b, *psi, *b_spe, *q_beam_j, *b_beam_j;
    double *cu, *dcu;   // J_perp predictor in place: dcu *= dxi, cu(1:2) += dcu
    double dxi;
    double *e_spe, *q_spe;
    double *h_e, *h_b, *h_psi, *h_b_spe, *h_e_spe, *h_cu, *h_q_spe;
    double *q_beam_f1, *b_beam_f1;
    IterState *it;
    const double *conv_blk;   // both parities
    int nblk;
    int iter_max;
};

// thread per (node i = 0..nr+1, mode array k)
__global__ __launch_bounds__(kNodeThreads) void k_epilogue(EpiArgs a)
{
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (blockIdx.x == 0 && k == 0 && threadIdx.x < 64) {
        // residuals of the last iteration executed (rel/abs of 'compare')
        const int iters = a.it->iters;
        double rel, ab;
        conv_residuals(a.conv_blk + (size_t)(iters & 1) * 2 * a.nblk, a.nblk, rel, ab);
        if (threadIdx.x == 0) {
            a.it->rel = rel;
            a.it->abs_ = ab;
            a.it->total_iters += iters;
            if (iters >= a.iter_max) a.it->at_max += 1;
        }
    }
    if (i > nr + 1) return;
    // every load first, then every store: the arrays are distinct, and
    // interleaved copies would wait for each load in turn (the compiler must
    // assume a store may alias the next load)
    double ve[3], vb[3], vcu[3], vbs[3], vbb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ve[c] = FX(a.e, c, k, i);
        vb[c] = FX(a.b, c, k, i);
        vcu[c] = FX(a.cu, c, k, i);
        vbs[c] = FX(a.b_spe, c, k, i);
        vbb[c] = FX(a.b_beam_j, c, k, i);
    }
    // E_perp of the plasma fields alone (e_spe, for the next slice's pipeline
    // and diagnostics)
    const bool inner = i >= 1 && i <= nr;
    double ev[2];
    if (inner) {
        et_node(g, vbs[0], vbs[1], FieldArr{a.psi, nm, ld}, k, i, ev);
    } else {
        ev[0] = FX(a.e_spe, 0, k, i);
        ev[1] = FX(a.e_spe, 1, k, i);
    }
    const double es2 = FX(a.e_spe, 2, k, i);
    const size_t o = (size_t)k * ld + i;
    const double qs = a.q_spe[o], ps = a.psi[o], qb = a.q_beam_j[o];
    const double dc0 = FX(a.dcu, 0, k, i), dc1 = FX(a.dcu, 1, k, i);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        FX(a.h_e, c, k, i) = ve[c];
        FX(a.h_b, c, k, i) = vb[c];
        FX(a.h_cu, c, k, i) = vcu[c];
        FX(a.h_b_spe, c, k, i) = vbs[c];
        FX(a.b_beam_f1, c, k, i) = vbb[c];
    }
    if (inner) {
        FX(a.e_spe, 0, k, i) = ev[0];
        FX(a.e_spe, 1, k, i) = ev[1];
    }
    FX(a.h_e_spe, 0, k, i) = ev[0];
    FX(a.h_e_spe, 1, k, i) = ev[1];
    FX(a.h_e_spe, 2, k, i) = es2;
    // q_spe = (rho - Jz) + Jz (species2d%cbq)
    const double q = qs + vcu[2];
    a.q_spe[o] = q;
    a.h_q_spe[o] = q;
    a.h_psi[o] = ps;
    a.q_beam_f1[o] = qb;
    // J_perp predictor for the next slice, after the history copy of cu
    // (simulation_class.f03:425-426)
    const double d0 = a.dxi * dc0, d1 = a.dxi * dc1;
    FX(a.dcu, 0, k, i) = d0;
    FX(a.dcu, 1, k, i) = d1;
    FX(a.cu, 0, k, i) = vcu[0] + d0;
    FX(a.cu, 1, k, i) = vcu[1] + d1;
}

int launch_epilogue(qpad_ctx *c, int j)
{
    const size_t s1 = (size_t)j * c->g.nm * c->g.ld, s3 = 3 * s1;
    EpiArgs a;
    a.g = field_common(c);
    a.e = c->f[QPAD_F_E];
    a.b = c->f[QPAD_F_B];
    a.psi = c->f[QPAD_F_PSI];
    a.cu = c->f[QPAD_F_CU];
    a.dcu = c->f[QPAD_F_DCU];
    a.dxi = c->g.dxi;
    a.b_spe = c->f[QPAD_F_B_SPE];
    a.q_beam_j = c->h_q_beam + s1;
    a.b_beam_j = c->h_b_beam + s3;
    a.e_spe = c->f[QPAD_F_E_SPE];
    a.q_spe = c->f[QPAD_F_Q_SPE];
    a.h_e = c->h_e + s3;
    a.h_b = c->h_b + s3;
    a.h_psi = c->h_psi + s1;
    a.h_b_spe = c->h_b_spe + s3;
    a.h_e_spe = c->h_e_spe + s3;
    a.h_cu = c->h_cu + s3;
    a.h_q_spe = c->h_q_spe + s1;
    a.q_beam_f1 = c->f[QPAD_F_Q_BEAM];
    a.b_beam_f1 = c->f[QPAD_F_B_BEAM];
    a.it = c->it;
    a.conv_blk = c->conv_blk;
    a.nblk = conv_blocks(c->g.nr);
    a.iter_max = c->g.iter_max;
    hipLaunchKernelGGL(k_epilogue, dim3((unsigned)node_blocks(c->g.nr), c->g.nm), dim3(kNodeThreads), 0,
                       c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// start of a slice (simulation_class.f03:342-377): rho - Jz of the plasma,
// psi, Bz and Ez from the J predictor, 'record', then E_perp and b
// ---------------------------------------------------------------------------
struct StartSrcArgs {
    FieldCommon g;
    const double *P0[kMaxSpecies], *P1[kMaxSpecies], *qn[kMaxSpecies];
    const double *O0[kMaxSpecies], *O1[kMaxSpecies];
    const int32_t *os[kMaxSpecies];
    const double *cu;
    double *q_spe, *src;
};

// thread per (node i = 0..nr+1, mode array k): q_spe = rho - Jz of every
// species from the charge-deposit partials (axis rules and 1/r of
// qdeposit_part2d, part2d_class.f03:313-336) plus the neutralising
// backgrounds; the psi, Bz and Ez right-hand sides at node i
template <int NS>
__global__ __launch_bounds__(kNodeThreads) void k_start_src(StartSrcArgs a)
{
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y, m = mode_of(k);
    if (i > nr + 1) return;
    const int cl0 = min(max(i - 1, 0), nr - 1), cl1 = min(max(i - 2, 0), nr - 1);
    const size_t c0 = (size_t)cl0 * nm + k;
    const size_t c1 = (size_t)cl1 * nm + k;
    // any overflow rows (crowded cells) among the workgroup's cells?
    // (uniform over the workgroup)
    bool ovf = false;
    {
        const int ib = blockIdx.x * blockDim.x;
        const int lo = max(ib - 2, 0), hi = min(ib + (int)blockDim.x - 2, nr - 1);
#pragma unroll
        for (int s = 0; s < NS; s++) ovf = ovf || a.os[s][hi + 1] != a.os[s][lo];
    }
    double d0[NS], d1[NS], qb[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        d0[s] = a.P0[s][c0];
        d1[s] = a.P1[s][c1];
        qb[s] = a.qn[s] ? a.qn[s][(size_t)k * ld + i] : 0.0;
    }
    if (ovf) {
        // overflow rows of crowded cells, in row order (cell_partial)
#pragma unroll
        for (int s = 0; s < NS; s++) {
            d0[s] = add_rows(d0[s], a.O0[s], a.os[s][cl0], a.os[s][cl0 + 1], nm, k);
            d1[s] = add_rows(d1[s], a.O1[s], a.os[s][cl1], a.os[s][cl1 + 1], nm, k);
        }
    }
    double v = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (i >= 1 && i <= nr) v += d0[s];
        if (i >= 2) v += d1[s];
    }
    if (i == 0) v = 0.0;
    else if (i == 1) v = (k == 0) ? 8.0 * v : 0.0;
    else v = v * (1.0 / (double)(i - 1));
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (a.qn[s]) v = v + qb[s];
    a.q_spe[(size_t)k * ld + i] = v;
    if (i < 1 || i > nr) return;
    const FieldArr cu{a.cu, nm, ld};
    a.src[(size_t)src_psi(M, nm, k) * nr + i - 1] = -1.0 * v;
    a.src[(size_t)src_bz(M, k) * nr + i - 1] = bz_source(g, cu, k, i);
    a.src[(size_t)src_ez(M, nm, k) * nr + i - 1] = (m == 0) ? ez_source_0(g, cu, i) : ez_source_m(g, cu, k, i);
}

struct StartSolveArgs {
    FieldCommon g;
    const double *src, *cu;
    double *psi, *b_spe, *e;
};

// workgroups: psi of each mode array, Bz of each, Ez of each (none of them
// reads B_perp, so they run before the iteration's seed is in place)
template <int K>
__global__ __launch_bounds__(kSolveThreads) void k_start_solve(StartSolveArgs a)
{
    extern __shared__ double stage[];
    __shared__ double red[kRed];
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int role = b / nm, k = b % nm, m = mode_of(k);
    const int row = role == 0 ? src_psi(M, nm, k) : (role == 1 ? src_bz(M, k) : src_ez(M, nm, k));
    double v[K], d[1][K];
    load_rows<K>(a.src + (size_t)row * nr, nr, v);
    if (role == 2 && m == 0) {
        // discrete Gauss law on the axis row of Ez
        double part = 0.0;
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            const int i = t + kk * kSolveThreads + 1;
            if (i >= 2 && i <= nr - 2) part += v[kk] * (double)(i - 1);
        }
        double div = block_sum(part, red);
        if (t == 0) {
            div = div - ez_axis_edge(g, FieldArr{a.cu, nm, ld});
            v[0] = -8.0 * div;
        }
    }
    stage_in<K>(stage, v, nr, d[0]);
    const FactorT f[1] = {dfactor(g, role == 0 ? FK_PSI : (role == 1 ? FK_BZ : FK_EZ), m)};
    seg_solve<K, 1>(f, d, red);
    stage_out<K>(stage, d[0], v);
#pragma unroll
    for (int kk = 0; kk < K; kk++) {
        const int r = t + kk * kSolveThreads, i = r + 1;
        if (r >= nr) continue;
        const double x = (m > 0 && i == 1) ? 0.0 : v[kk];
        if (role == 0) a.psi[(size_t)k * ld + i] = x;
        else if (role == 1) FX(a.b_spe, 2, k, i) = x;
        else FX(a.e, 2, k, i) = x;
    }
}

struct EperpArgs {
    FieldCommon g;
    double *b_spe;
    const double *b_beam, *psi;
    double *b_out, *e;
    // learned seed of the iteration (B_perp components 0, 1 of nodes
    // 1..nr), copied into b_spe here, or null
    const double *seed;
    double *conv_re, *conv_im;
    IterState *it;
};

// thread per (node i = 0..nr+1, mode array k): the iteration's seed into
// b_spe (when given), b = b_spe + b_beam and E_perp from b and psi
// (field_e_class.f03:443-504); the k = 0 threads also run 'record' of
// sum_m |B_phi,m| at their node (simulation_class.f03:549-558) and reset
// the iteration state.  The first kernel of the slice that reads B_perp.
__global__ __launch_bounds__(kNodeThreads) void k_eperp(EperpArgs a)
{
    const FieldCommon &g = a.g;
    const int nm = g.nm, ld = g.ld, nr = g.nr, M = g.M;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i > nr + 1) return;
    const bool inner = i >= 1 && i <= nr;
    // B_perp of the iteration's start: the seed, or b_spe as it is
    const double *bperp = (a.seed && inner) ? a.seed : a.b_spe;
    if (k == 0) {
        if (i == 0) {
            a.it->done = 0;
            a.it->iters = 0;
            a.it->counter = 0;
        }
        if (inner) {
            double re = 0.0, im = 0.0;
            for (int mm = 0; mm <= M; mm++) {
                re = re + fabs(FX(bperp, 1, kre(mm), i));
                if (mm > 0) im = im + fabs(FX(bperp, 1, kim(mm), i));
            }
            a.conv_re[i] = re;
            a.conv_im[i] = im;
        }
    }
    double bv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double bs = FX(c < 2 ? bperp : a.b_spe, c, k, i);
        if (c < 2 && a.seed && inner) FX(a.b_spe, c, k, i) = bs;
        bv[c] = bs + FX(a.b_beam, c, k, i);
        FX(a.b_out, c, k, i) = bv[c];
    }
    if (!inner) return;
    double ev[2];
    et_node(g, bv[0], bv[1], FieldArr{a.psi, nm, ld}, k, i, ev);
    FX(a.e, 0, k, i) = ev[0];
    FX(a.e, 1, k, i) = ev[1];
}

int src_rows_host(int M, int nm) { return src_rows(M, nm); }

// sources of the slice start (q_spe, the psi / Bz / Ez right-hand sides)
int launch_slice_sources(qpad_ctx *c)
{
    const FieldCommon g = field_common(c);
    const int nm = c->g.nm;
    const dim3 ngrid((unsigned)node_blocks(c->g.nr), nm);
    StartSrcArgs s;
    memset(&s, 0, sizeof(s));
    s.g = g;
    for (int q = 0; q < c->nspecies; q++) {
        s.P0[q] = c->sp[q].P0;
        s.P1[q] = c->sp[q].P1;
        // the neutralising background, or a neutral gas's ion charge
        s.qn[q] = (c->sp[q].cfg.neutralized || c->sp[q].neutral) ? c->sp[q].qn : nullptr;
        s.O0[q] = c->sp[q].O0;
        s.O1[q] = c->sp[q].O1;
        s.os[q] = c->sp[q].ov_start;
    }
    s.cu = c->f[QPAD_F_CU];
    s.q_spe = c->f[QPAD_F_Q_SPE];
    s.src = c->src;
    switch (c->nspecies) {
    case 1: hipLaunchKernelGGL(k_start_src<1>, ngrid, dim3(kNodeThreads), 0, c->stream, s); break;
    case 2: hipLaunchKernelGGL(k_start_src<2>, ngrid, dim3(kNodeThreads), 0, c->stream, s); break;
    case 3: hipLaunchKernelGGL(k_start_src<3>, ngrid, dim3(kNodeThreads), 0, c->stream, s); break;
    default: hipLaunchKernelGGL(k_start_src<4>, ngrid, dim3(kNodeThreads), 0, c->stream, s); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// psi, Bz from the J predictor and Ez of the slice start (no B_perp read)
int launch_start_solves(qpad_ctx *c)
{
    const FieldCommon g = field_common(c);
    const int nm = c->g.nm;
    StartSolveArgs a;
    a.g = g;
    a.src = c->src;
    a.cu = c->f[QPAD_F_CU];
    a.psi = c->f[QPAD_F_PSI];
    a.b_spe = c->f[QPAD_F_B_SPE];
    a.e = c->f[QPAD_F_E];
    QPAD_DISPATCH_K(c->g.nr, k_start_solve, dim3(3 * nm), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// the iteration's seed into b_spe (seed: the learned seed or null), the
// first 'record', b and E_perp for the first deposit
int launch_start_eperp(qpad_ctx *c, int j, const double *seed)
{
    const int nm = c->g.nm;
    const dim3 ngrid((unsigned)node_blocks(c->g.nr), nm);
    EperpArgs e;
    e.g = field_common(c);
    e.b_spe = c->f[QPAD_F_B_SPE];
    e.b_beam = c->h_b_beam + (size_t)j * 3 * nm * c->g.ld;
    e.psi = c->f[QPAD_F_PSI];
    e.b_out = c->f[QPAD_F_B];
    e.e = c->f[QPAD_F_E];
    e.seed = seed;
    e.conv_re = c->conv_re;
    e.conv_im = c->conv_im;
    e.it = c->it;
    hipLaunchKernelGGL(k_eperp, ngrid, dim3(kNodeThreads), 0, c->stream, e);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int K>
static int set_lds_attr_slice()
{
    // the solve stage; it later holds one psi row (nr + 2 <= stage)
    const int bytes = 8 * stage_doubles<K>();
    QPAD_HIP_CHECK(hipFuncSetAttribute((const void *)k_solve<K>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    QPAD_HIP_CHECK(hipFuncSetAttribute((const void *)k_start_solve<K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       bytes));
    return 0;
}

int slice_kernels_init()
{
    const void *srcs[] = {(const void *)k_src<1>, (const void *)k_src<2>, (const void *)k_src<3>,
                          (const void *)k_src<4>};
    for (const void *k : srcs)
        QPAD_HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(sizeof(double) * 8 * kMaxNm * kSrcTL)));
    QPAD_TRY(set_lds_attr_slice<1>());
    QPAD_TRY(set_lds_attr_slice<2>());
    QPAD_TRY(set_lds_attr_slice<4>());
    QPAD_TRY(set_lds_attr_slice<8>());
    QPAD_TRY(set_lds_attr_slice<16>());
    return 0;
}

}  // namespace qpad
