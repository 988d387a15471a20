// k_plasma.hip -- 2D plasma species on the GPU (FP64): lattice injection,
// radial-cell sort, rho - Jz deposit, the fused gather + robust predictor +
// current/acceleration/momentum-flux deposit, and the fused push_u + push_x.
//
// Particles are kept sorted by radial cell (cell c = int(r/dr), nodes c+1 and
// c+2).  Deposits run one wavefront per chunk of at most kDepChunk particles
// of a cell: each lane accumulates its particles in registers, the wavefront
// reduces with a fixed butterfly and writes the chunk's partials for node c+1
// (P0) and node c+2 (P1) once.  A cell's first chunk goes to P0/P1[c], its
// further chunks (crowded cells only) to the overflow rows O0/O1 the sort
// lists; the field kernels add P0[i-1] + P1[i-2] per node, each with its
// cell's overflow rows in row order.  No atomics, deterministic order.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "kernels.h"
#include "sort_config.h"
#include "rng.h"
#include "wave_reduce.h"
#include "conv.h"

namespace qpad {

#define FX(p, c, k, i) (p)[((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)]

// Radius and normalised position without contraction, so that every kernel
// computes the same cell for the same particle.
__device__ inline double radius(double x1, double x2)
{
    return sqrt(__dadd_rn(__dmul_rn(x1, x1), __dmul_rn(x2, x2)));
}

__device__ inline uint32_t cell_key(double x1, double x2, double idr, double edge, int nr)
{
    double r = radius(x1, x2);
    if (!(r < edge)) return (uint32_t)nr;   // leaves the box (update_bound)
    int c = (int)__dmul_rn(r, idr);
    return (uint32_t)(c < nr ? c : nr - 1);
}

// ---------------------------------------------------------------------------
// injection: fdist2d_class.f03:289-357 (lattice theta).  A uniform plasma
// keeps every lattice point; a profiled one (den_perp given) keeps the points
// whose den_lon * den_perp * density reaches den_min, at the exclusive prefix
// `off` of the kept flags, so that the kept particles keep QPAD's order and
// their index (which also draws their thermal momenta) counts kept points.
// ---------------------------------------------------------------------------
struct InjArgs {
    int nr, ppc1, ppc2, nth;
    int64_t n;
    double dr, idr, edge, density, coef, qm, dtheta, den_lon, den_min;
    double uth0, uth1, uth2;
    uint64_t seed;
    const double *den;       // transverse density per lattice point (null: 1)
    const int32_t *off;      // kept points' exclusive prefix (null: all kept)
    double *x1, *x2, *p1, *p2, *p3, *q, *gamma, *psi;
    int32_t *id;
    uint32_t *key;
};

// kept flag of every lattice point, 0 in the pad entry n
__global__ void k_inject_keep(int64_t n, const double *den, double den_lon, double density, double den_min,
                              int32_t *keep)
{
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l > n) return;
    keep[l] = l < n ? (den_lon * den[l] * density < den_min ? 0 : 1) : 0;
}

__global__ void k_inject(InjArgs a)
{
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.n) return;
    int64_t ip = l;
    if (a.off) {
        if (a.off[l + 1] == a.off[l]) return;
        ip = a.off[l];
    }
    int64_t t = l;
    const int i2 = (int)(t % a.ppc2); t /= a.ppc2;
    const int i1 = (int)(t % a.ppc1); t /= a.ppc1;
    const int i = (int)(t % a.nr);
    const int j = (int)(t / a.nr);
    // 1-based loop indices of the reference: j+1, i+1, i1+1, i2+1
    double rn = ((double)(i1 + 1) - 0.5) / (double)a.ppc1 + (double)i;
    double theta = (((double)(i2 + 1) - 0.5) / (double)a.ppc2 + (double)(j + 1) - 1.0) * a.dtheta;
    double x1 = rn * a.dr * cos(theta), x2 = rn * a.dr * sin(theta);
    double th[4] = {0.0, 0.0, 0.0, 0.0};
    if (a.uth0 != 0.0 || a.uth1 != 0.0 || a.uth2 != 0.0) normal4(a.seed, (uint64_t)ip, 0u, th);
    double p1 = a.uth0 * th[0], p2 = a.uth1 * th[1], p3 = a.uth2 * th[2];
    double gam = sqrt(1.0 + p1 * p1 + p2 * p2 + p3 * p3);
    const double den_perp = a.den ? a.den[l] : 1.0;
    a.x1[ip] = x1;
    a.x2[ip] = x2;
    a.q[ip] = rn * den_perp * a.den_lon * a.density * a.coef;
    a.p1[ip] = p1;
    a.p2[ip] = p2;
    a.p3[ip] = p3;
    a.gamma[ip] = gam;
    a.psi[ip] = (1.0 - gam + p3) / a.qm;
    a.id[ip] = (int32_t)ip;
    a.key[ip] = cell_key(x1, x2, a.idr, a.edge, a.nr);
}

int launch_inject(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    const qpad_species_cfg &cf = sp.cfg;
    InjArgs a;
    a.nr = c->g.nr;
    a.ppc1 = cf.ppc1;
    a.ppc2 = cf.ppc2;
    a.nth = cf.num_theta;
    a.n = (int64_t)cf.num_theta * c->g.nr * cf.ppc1 * cf.ppc2;
    a.dr = c->g.dr;
    a.idr = 1.0 / c->g.dr;
    a.edge = (double)c->g.nr * c->g.dr;
    a.density = cf.density;
    a.coef = copysign(1.0, cf.qm) / ((double)(cf.ppc1 * cf.ppc2) * (double)cf.num_theta);
    a.qm = cf.qm;
    a.dtheta = 2.0 * (4.0 * atan(1.0)) / (double)cf.num_theta;
    a.den_lon = sp.den_perp ? sp.den_lon : 1.0;
    a.den_min = cf.den_min;
    a.uth0 = cf.uth[0];
    a.uth1 = cf.uth[1];
    a.uth2 = cf.uth[2];
    a.seed = cf.seed;
    a.den = sp.den_perp;
    a.off = nullptr;
    a.x1 = sp.x1; a.x2 = sp.x2; a.p1 = sp.p1; a.p2 = sp.p2; a.p3 = sp.p3;
    a.q = sp.q; a.gamma = sp.gamma; a.psi = sp.psi; a.id = sp.id;
    a.key = sp.key;
    int64_t kept = a.n;
    if (sp.den_perp) {
        hipLaunchKernelGGL(k_inject_keep, dim3((unsigned)((a.n + 1 + 255) / 256)), dim3(256), 0, c->stream, a.n,
                           sp.den_perp, a.den_lon, a.density, a.den_min, sp.inj_keep);
        QPAD_HIP_CHECK(hipGetLastError());
        size_t bytes = sp.inj_tmp_bytes;
        QPAD_HIP_CHECK(rocprim::exclusive_scan(sp.inj_tmp, bytes, sp.inj_keep, sp.inj_off, 0, (size_t)a.n + 1,
                                               rocprim::plus<int32_t>(), c->stream));
        QPAD_HIP_CHECK(hipMemcpyAsync(sp.inj_host, sp.inj_off + a.n, sizeof(int32_t), hipMemcpyDeviceToHost,
                                      c->stream));
        QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
        kept = *sp.inj_host;
        a.off = sp.inj_off;
    } else if (1.0 * 1.0 * cf.density < cf.den_min) {
        // den_lon * den_perp * density < den_min drops every particle
        kept = 0;
    }
    sp.npp = kept;
    if (kept > 0) {
        hipLaunchKernelGGL(k_inject, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, c->stream, a);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// scan scratch of the profiled injection of n lattice points
int inject_scan_bytes(int64_t n, size_t *bytes)
{
    int32_t *z = nullptr;
    QPAD_HIP_CHECK(rocprim::exclusive_scan(nullptr, *bytes, z, z, 0, (size_t)n + 1, rocprim::plus<int32_t>()));
    return 0;
}

// ---------------------------------------------------------------------------
// cell start offsets of sorted keys (the beam's (xi, r)-cell sort)
// ---------------------------------------------------------------------------
// cs[c] = first position p with key[p] >= c (lower bound in the sorted
// keys), for c = 0 .. ncs-1: one thread per cell, so empty runs of cells
// cost nothing extra
__global__ __launch_bounds__(256) void k_cellstart(const uint32_t *key, int64_t n, int ncs, int32_t *cs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncs) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (key[mid] < (uint32_t)c) lo = mid + 1;
        else hi = mid;
    }
    cs[c] = (int32_t)lo;
}

int launch_cell_start(qpad_ctx *c, const uint32_t *key, int64_t n, int ncs, int32_t *cs)
{
    hipLaunchKernelGGL(k_cellstart, dim3((unsigned)((ncs + 255) / 256)), dim3(256), 0, c->stream, key, n, ncs,
                       cs);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// scratch of rocPRIM's radix sort (uint32 keys and values) and exclusive scan
int sort_scratch(size_t n, int bits, size_t *bytes)
{
    size_t b1 = 0, b2 = 0;
    uint32_t *z = nullptr;
    QPAD_HIP_CHECK(rocprim::radix_sort_pairs<SortConfig>(nullptr, b1, z, z, z, z, n, 0, bits));
    QPAD_HIP_CHECK(rocprim::exclusive_scan(nullptr, b2, z, z, 0u, n, rocprim::plus<uint32_t>()));
    *bytes = std::max(b1, b2) + 256;
    return 0;
}

__global__ void k_keys(const double *x1, const double *x2, uint32_t *key, int64_t n, double idr, double edge,
                       int nr)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    key[p] = cell_key(x1[p], x2[p], idr, edge, nr);
}

int launch_keys(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    if (sp.npp == 0) return 0;
    hipLaunchKernelGGL(k_keys, dim3((unsigned)((sp.npp + 255) / 256)), dim3(256), 0, c->stream, sp.x1, sp.x2,
                       sp.key, sp.npp, 1.0 / c->g.dr, (double)c->g.nr * c->g.dr, c->g.nr);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// rho - Jz deposit: qdeposit_part2d (part2d_class.f03:231-359), per-cell
// partials; axis rules and 1/r are applied by the consumer
// ---------------------------------------------------------------------------
struct QDepArgs {
    int nr;
    double idr;
    const double *x1, *x2, *q;
    // susceptibility deposit (CHI): the particle's psi and q/m
    const double *psi;
    double qbm;
    const int32_t *cs;
    double *P0, *P1;
    // overflow chunks (listed by the sort) and their partial rows
    const int32_t *os;
    const int4 *desc;
    double *O0, *O1;
    int n_ov_max;
};

// One wave per chunk of at most kDepChunk particles of one cell.  The first
// ceil(nr / CPW) workgroups take the first chunk of cell (workgroup * CPW +
// wave) into P0/P1; the workgroups after them take the overflow chunks
// (row = (workgroup - ceil(nr / CPW)) * CPW + wave) into O0/O1.
// CPW: waves per workgroup, a launch shape the tuner chooses
// (qpad_tune_launch); every wave's work and result are the same for every
// shape.
struct DepChunk {
    int cell, pb, pe, row;
    bool ovf, live;
};

template <int CPW>
__device__ __forceinline__ DepChunk dep_chunk(int nr, const int32_t *cs, const int32_t *os, const int4 *desc,
                                              int n_ov_max, int wave)
{
    DepChunk k;
    const int ncg = (nr + CPW - 1) / CPW;
    k.ovf = (int)blockIdx.x >= ncg;
    if (!k.ovf) {
        const int c = blockIdx.x * CPW + wave;
        k.live = c < nr;
        k.cell = min(c, nr - 1);
        k.row = k.cell;
        const int p0 = cs[k.cell], p1 = cs[k.cell + 1];
        k.pb = p0;
        k.pe = min(p1, p0 + kDepChunk);
    } else {
        const int o = ((int)blockIdx.x - ncg) * CPW + wave;
        const int nov = os[nr];
        const int4 d = desc[min(o, n_ov_max - 1)];
        k.live = o < nov;
        k.cell = min(max(d.x, 0), nr - 1);
        k.row = min(o, n_ov_max - 1);
        k.pb = d.y;
        k.pe = k.live ? d.z : d.y;
    }
    return k;
}

// overflow workgroups none of whose waves has a chunk (the grid is sized
// for the most chunks the particle count allows); uniform over the workgroup
template <int CPW>
__device__ __forceinline__ bool dep_group_idle(int nr, const int32_t *os)
{
    const int ncg = (nr + CPW - 1) / CPW;
    return (int)blockIdx.x >= ncg && ((int)blockIdx.x - ncg) * CPW >= os[nr];
}

// overflow rows the grid provides for (at most) npp particles in the box
inline int dep_grid(int nr, int64_t npp, int n_ov_max, int cpw)
{
    const int nov = (int)std::min<int64_t>(npp / kDepChunk + 1, n_ov_max);
    return (nr + cpw - 1) / cpw + (nov + cpw - 1) / cpw;
}

// the arguments of every species of one launch (workgroup rows by species)
struct QDepMulti {
    QDepArgs s[kMaxSpecies];
};

// CHI: deposit_chi_part2d (part2d_class.f03:361-476) instead: each particle
// deposits -(q/m) q / (1 - (q/m) psi) in place of q
template <int M, int CPW, bool CHI = false>
__global__ __launch_bounds__(64 * CPW) void k_qdeposit(const QDepMulti m)
{
    const QDepArgs a = m.s[__builtin_amdgcn_readfirstlane(blockIdx.y)];
    constexpr int NM = 2 * M + 1;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const DepChunk ch = dep_chunk<CPW>(a.nr, a.cs, a.os, a.desc, a.n_ov_max, wave);
    if (!ch.live) return;
    const int c = ch.cell, p0 = ch.pb, p1 = ch.pe;
    // acc[k]: node ix, acc[NM + k]: node ix + 1; padded to a multiple of 16
    constexpr int NB = (2 * NM + 15) / 16;
    double acc[16 * NB];
#pragma unroll
    for (int v = 0; v < 16 * NB; v++) acc[v] = 0.0;
    // wave-uniform passes, software-pipelined as in k_amjdeposit; lanes past
    // the chunk's end deposit zero
    const int npass = __builtin_amdgcn_readfirstlane((p1 - p0 + 63) >> 6);
    const int pl = max(p1 - 1, p0);
    int p = p0 + lane;
    int pc = min(p, pl);
    double nx1 = a.x1[pc], nx2 = a.x2[pc], nq = a.q[pc];
    double npsi = CHI ? a.psi[pc] : 0.0;
    for (int k = 0; k < npass; k++, p += 64) {
        const double x1 = nx1, x2 = nx2;
        double q = p < p1 ? nq : 0.0;
        if constexpr (CHI) q = p < p1 ? -1.0 * a.qbm * nq / (1.0 - a.qbm * npsi) : 0.0;
        pc = min(p + 64, pl);
        nx1 = a.x1[pc];
        nx2 = a.x2[pc];
        nq = a.q[pc];
        if constexpr (CHI) npsi = a.psi[pc];
        double pos = __dmul_rn(radius(x1, x2), a.idr);
        double ph0r = x1 / pos * a.idr, ph0i = -x2 / pos * a.idr;
        double w1 = pos - (double)c, w0 = 1.0 - w1;
        double phr = 1.0 * q, phi = 0.0 * q;
        acc[0] += w0 * phr;
        acc[NM] += w1 * phr;
#pragma unroll
        for (int m = 1; m <= M; m++) {
            double tr = phr * ph0r - phi * ph0i, ti = phr * ph0i + phi * ph0r;
            phr = tr;
            phi = ti;
            acc[2 * m - 1] += w0 * phr;
            acc[NM + 2 * m - 1] += w1 * phr;
            acc[2 * m] += w0 * phi;
            acc[NM + 2 * m] += w1 * phi;
        }
    }
    wave_reduce_scatter<NB>(acc);
    const int g = lane >> 2;
    double *D0 = ch.ovf ? a.O0 : a.P0, *D1 = ch.ovf ? a.O1 : a.P1;
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int v = g * NB + i;
        if ((lane & 3) == (i & 3) && v < 2 * NM) {
            if (v < NM) D0[(size_t)ch.row * NM + v] = acc[i];
            else D1[(size_t)ch.row * NM + v - NM] = acc[i];
        }
    }
}

static QDepArgs qdep_args(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    QDepArgs a;
    a.nr = c->g.nr;
    a.idr = 1.0 / c->g.dr;
    a.x1 = sp.x1; a.x2 = sp.x2; a.q = sp.q;
    a.psi = sp.psi;
    a.qbm = sp.qbm;
    a.cs = sp.cell_start;
    a.P0 = sp.P0;
    a.P1 = sp.P1;
    a.os = sp.ov_start;
    a.desc = sp.ov_desc;
    a.O0 = sp.O0;
    a.O1 = sp.O1;
    a.n_ov_max = sp.n_ov_max;
    return a;
}

int launch_qdeposit(qpad_ctx *c, int s) { return launch_qdeposit_range(c, s, s + 1); }

int launch_qdeposit_all(qpad_ctx *c) { return c->nspecies ? launch_qdeposit_range(c, 0, c->nspecies) : 0; }

// the charge deposit of the species in m.s[0 .. ns-1], gx workgroups per
// species (chi: the susceptibility deposit)
static int qdeposit_launch(qpad_ctx *c, const QDepMulti &m, int ns, int gx, bool chi = false)
{
    const int cpw = c->dep_cells;
    dim3 grid((unsigned)gx, (unsigned)ns), blk(64 * cpw);
#define QPAD_QDEP(MM, CC)                                                                        \
    do {                                                                                         \
        if (chi) hipLaunchKernelGGL((k_qdeposit<MM, CC, true>), grid, blk, 0, c->stream, m);     \
        else hipLaunchKernelGGL((k_qdeposit<MM, CC, false>), grid, blk, 0, c->stream, m);        \
    } while (0)
#define QPAD_QDEP_M(MM)                              \
    do {                                             \
        if (cpw == 2) QPAD_QDEP(MM, 2);              \
        else if (cpw == 8) QPAD_QDEP(MM, 8);         \
        else QPAD_QDEP(MM, 4);                       \
    } while (0)
    switch (c->g.M) {
    case 0: QPAD_QDEP_M(0); break;
    case 1: QPAD_QDEP_M(1); break;
    case 2: QPAD_QDEP_M(2); break;
    case 3: QPAD_QDEP_M(3); break;
    default: QPAD_QDEP_M(4); break;
    }
#undef QPAD_QDEP_M
#undef QPAD_QDEP
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// species s0 .. s1-1 in one launch
int launch_qdeposit_range(qpad_ctx *c, int s0, int s1)
{
    QDepMulti m;
    memset(&m, 0, sizeof(m));
    int gx = 1;
    for (int s = s0; s < s1; s++) {
        m.s[s - s0] = qdep_args(c, s);
        gx = std::max(gx, dep_grid(c->g.nr, c->sp[s].npp, c->sp[s].n_ov_max, c->dep_cells));
    }
    return qdeposit_launch(c, m, s1 - s0, gx);
}

// species s with its partial rows in caller-given memory instead of the
// species' own: rows.P0/P1 hold nr rows and rows.O0/O1 n_ov_max rows of nm
// values each.  Every row a consumer reads is written (empty cells write
// zeros), so the rows need no clearing.
int launch_qdeposit_rows(qpad_ctx *c, int s, const DepRows &rows)
{
    QDepMulti m;
    memset(&m, 0, sizeof(m));
    m.s[0] = qdep_args(c, s);
    m.s[0].P0 = rows.P0;
    m.s[0].P1 = rows.P1;
    m.s[0].O0 = rows.O0;
    m.s[0].O1 = rows.O1;
    const int gx = std::max(1, dep_grid(c->g.nr, c->sp[s].npp, c->sp[s].n_ov_max, c->dep_cells));
    return qdeposit_launch(c, m, 1, gx);
}

// the susceptibility deposit of species s into caller-given rows (as
// launch_qdeposit_rows)
int launch_chi_deposit_rows(qpad_ctx *c, int s, const DepRows &rows)
{
    QDepMulti m;
    memset(&m, 0, sizeof(m));
    m.s[0] = qdep_args(c, s);
    m.s[0].P0 = rows.P0;
    m.s[0].P1 = rows.P1;
    m.s[0].O0 = rows.O0;
    m.s[0].O1 = rows.O1;
    const int gx = std::max(1, dep_grid(c->g.nr, c->sp[s].npp, c->sp[s].n_ov_max, c->dep_cells));
    return qdeposit_launch(c, m, 1, gx, true);
}

// one species' q (axis rules, 1/r) times `scale` into out (1 x nm x ld)
__global__ void k_combine_q(const double *P0, const double *P1, const double *O0, const double *O1,
                            const int32_t *os, double *out, int nr, int nm, int ld, double scale, int accumulate)
{
    int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nm * (nr + 2)) return;
    int i = (int)(tid % (nr + 2)), k = (int)(tid / (nr + 2));
    double v = 0.0;
    if (i >= 1 && i <= nr) v += cell_partial(P0, O0, os, i - 1, nm, k);
    if (i >= 2) v += cell_partial(P1, O1, os, i - 2, nm, k);
    if (i == 0) v = 0.0;
    else if (i == 1) v = (k == 0) ? 8.0 * v : 0.0;
    else v = v * (1.0 / (double)(i - 1));
    double *o = out + (size_t)k * ld + i;
    *o = accumulate ? *o + scale * v : scale * v;
}

int launch_combine_q_rows(qpad_ctx *c, int s, const DepRows &rows, double *out, double scale, bool accumulate)
{
    int64_t n = (int64_t)c->g.nm * (c->g.nr + 2);
    hipLaunchKernelGGL(k_combine_q, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, rows.P0, rows.P1,
                       rows.O0, rows.O1, c->sp[s].ov_start, out, c->g.nr, c->g.nm, c->g.ld, scale,
                       accumulate ? 1 : 0);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_combine_q(qpad_ctx *c, int s, double *out, double scale, bool accumulate)
{
    const Species &sp = c->sp[s];
    return launch_combine_q_rows(c, s, DepRows{sp.P0, sp.P1, sp.O0, sp.O1}, out, scale, accumulate);
}

// ---------------------------------------------------------------------------
// gather of a 3-component field at nodes ix, ix+1 (interp_part2d.f03:66-110)
// ---------------------------------------------------------------------------
// f(cc, k, n): component cc of mode array k at node ix + n (n = 0, 1); NC
// components (1: interp_field_part2d_scalar, :110-153)
template <int M, int NC, class F>
__device__ __forceinline__ void gather_n(const F &f, double w0, double w1, double pc, double ps, double out[NC])
{
#pragma unroll
    for (int cc = 0; cc < NC; cc++) {
        double v = 0.0;
        v = v + f(cc, 0, 0) * w0;
        v = v + f(cc, 0, 1) * w1;
        out[cc] = v;
    }
    double phr = 1.0, phi = 0.0;
#pragma unroll
    for (int m = 1; m <= M; m++) {
        double tr = phr * pc - phi * ps, ti = phr * ps + phi * pc;
        phr = tr;
        phi = ti;
        const double ph_re = 2.0 * phr, ph_im = 2.0 * phi;
#pragma unroll
        for (int cc = 0; cc < NC; cc++) {
            out[cc] = out[cc] + (f(cc, 2 * m - 1, 0) * ph_re - f(cc, 2 * m, 0) * ph_im) * w0;
            out[cc] = out[cc] + (f(cc, 2 * m - 1, 1) * ph_re - f(cc, 2 * m, 1) * ph_im) * w1;
        }
    }
}

template <int M, class F>
__device__ __forceinline__ void gather3(const F &f, double w0, double w1, double pc, double ps, double out[3])
{
    gather_n<M, 3>(f, w0, w1, pc, ps, out);
}

// nodes ix, ix + 1 of a slice field in HBM
struct NodePair {
    const double *f;
    int nm, ld, ix;
    __device__ __forceinline__ double operator()(int cc, int k, int n) const
    {
        return f[((size_t)cc * nm + k) * ld + ix + n];
    }
};

// nodes ix, ix + 1 of a slice field staged in LDS as [n][cc][k] (NC
// components)
template <int NM, int NC = 3>
struct NodePairLds {
    const double *s;
    __device__ __forceinline__ double operator()(int cc, int k, int n) const { return s[(n * NC + cc) * NM + k]; }
};

// the laser envelope (ar, ai) and its gradients (ar_grad, ai_grad: r, phi,
// z) at a particle, from nodes f(field, cc, k, n) (interp_field calls of
// part2d_class.f03:1100-1105, :1406-1409, :1996-2001)
struct LaserAt {
    double apr, api, aprg[3], apig[3];
};

template <int M, class FR, class FI, class GR, class GI>
__device__ __forceinline__ LaserAt gather_laser(const FR &far, const FI &fai, const GR &gar, const GI &gai, double w0,
                                               double w1, double pc, double ps)
{
    LaserAt l;
    gather_n<M, 1>(far, w0, w1, pc, ps, &l.apr);
    gather_n<M, 1>(fai, w0, w1, pc, ps, &l.api);
    gather_n<M, 3>(gar, w0, w1, pc, ps, l.aprg);
    gather_n<M, 3>(gai, w0, w1, pc, ps, l.apig);
    return l;
}

// ---------------------------------------------------------------------------
// amjdeposit_robust_part2d (part2d_class.f03:746-1010): gather E, B; robust
// predictor on a copy of the momentum; deposit cu (3), acu (2), amu (3) per
// mode.  Value index v = comp * NM + k, comp 0..2 cu, 3..4 acu, 5..7 amu.
// ---------------------------------------------------------------------------
struct JDepArgs {
    int nr, ld, nm;
    double idr, qbm, qtmh, idt;
    const double *x1, *x2, *p1, *p2, *p3, *q;
    double *gamma, *psi;
    int store_diag;
    const int32_t *cs;
    const double *e, *b;
    // the laser envelope and its gradients of the slice (PGC pushers)
    const double *ar, *ai, *arg, *aig;
    double *P0, *P1;
    IterState *it;
    // SPEC launches: convergence of the previous iteration from the field
    // kernel's per-workgroup maxima
    const double *conv_prev;
    int nblk;
    double reltol, abstol;
    // host-visible copy of the decision (done, iterations so far) written by
    // workgroup 0 of a SPEC launch, or null
    IterState *report;
    // overflow chunks (listed by the sort) and their partial rows
    const int32_t *os;
    const int4 *desc;
    double *O0, *O1;
    int n_ov_max;
};

// What one particle deposits (robust predictor, part2d_class.f03:877-960):
// the eight deposited quantities val[comp] (cu 0..2, dcu 3..4, amu 5..7),
// the linear weights w[n] of nodes ix + n and the mode phases of mode array
// k times q / (gamma - u_z), ph[k].  Value v = n * 8NM + comp * NM + k of
// the cell's deposit is the sum over its particles of
// w[n] ph[k] val[comp].
template <int M>
struct PartTerms {
    double val[8];
    double w[2];
    double ph[2 * M + 1];
};

// the products the reduce-scatter form accumulates: value v is
// (w[n] ph[k]) val[comp]
template <int M>
struct DepTerms {
    double val[8];
    double wph[2][2 * M + 1];
    __device__ __forceinline__ explicit DepTerms(const PartTerms<M> &t)
    {
#pragma unroll
        for (int c = 0; c < 8; c++) val[c] = t.val[c];
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int k = 0; k < 2 * M + 1; k++) wph[n][k] = t.w[n] * t.ph[k];
    }
    __device__ __forceinline__ double operator()(int v) const
    {
        constexpr int NM = 2 * M + 1;
        const int n = v / (8 * NM), r = v % (8 * NM);
        return wph[n][r % NM] * val[r / NM];
    }
};

// the values a deposit workgroup stages in LDS per wave: E, B ([n][cc][k]),
// then for the PGC pushers ar, ai ([n][k]) and ar_grad, ai_grad ([n][cc][k])
template <int NM, bool PGC>
constexpr int dep_staged() { return PGC ? 28 * NM : 12 * NM; }

// staged value l of the PGC layout of the cell whose left node is ix
template <int NM>
__device__ __forceinline__ double dep_stage_load_pgc(const double *e, const double *b, const double *ar,
                                                     const double *ai, const double *arg, const double *aig, int nm,
                                                     int ld, int l, int ix)
{
    const double *f;
    int nc, base;
    if (l < 12 * NM) {
        f = l < 6 * NM ? e : b;
        nc = 3;
        base = l < 6 * NM ? 0 : 6 * NM;
    } else if (l < 16 * NM) {
        f = l < 14 * NM ? ar : ai;
        nc = 1;
        base = l < 14 * NM ? 12 * NM : 14 * NM;
    } else {
        f = l < 22 * NM ? arg : aig;
        nc = 3;
        base = l < 22 * NM ? 16 * NM : 22 * NM;
    }
    const int r = l - base, n = r / (nc * NM), comp = (r / NM) % nc, k = r % NM;
    return f[((size_t)comp * nm + k) * ld + ix + n];
}

// gather and predictor of particle p of cell c; lanes past the chunk's end
// (valid false) deposit zero and store nothing.  STD: the standard pusher
// (amjdeposit_std_part2d, part2d_class.f03:478-744), which normalises by the
// particle's interpolated psi (psip) where the robust one uses gamma - u_z,
// and always stores the time-centred gamma.  PGC: the laser pushers
// (amjdeposit_std_pgc_part2d :1012-1308, amjdeposit_robust_pgc_part2d
// :1310-1602): every gamma carries gam_corr = (q/m)^2 |a|^2 / 2, the
// electric field the ponderomotive term -+ (q/m)/(2 gamma) a grad a
// (r, phi minus, z plus, with the initial gamma), and the robust form
// normalises its second half kick by gamma / (gamma - u_z) of the rotated
// momentum.  sf: the wave's staged values (dep_staged layout).
template <int M, bool STD, bool PGC>
__device__ __forceinline__ void particle_terms(const JDepArgs &a, const double *sf, int c, int p, bool valid,
                                               double x1, double x2, double pp1, double pp2, double pp3,
                                               double qp, double psip, PartTerms<M> &t)
{
    constexpr int NM = 2 * M + 1;
    double r = radius(x1, x2);
    const double pc = x1 / r, ps = x2 / r;
    double pos = __dmul_rn(r, a.idr);
    const double w1 = pos - (double)c, w0 = 1.0 - w1;
    double ep[3], bp[3];
    gather3<M>(NodePairLds<NM>{sf}, w0, w1, pc, ps, ep);
    gather3<M>(NodePairLds<NM>{sf + 6 * NM}, w0, w1, pc, ps, bp);
    double gam_corr = 0.0;
    LaserAt la;
    if constexpr (PGC) {
        la = gather_laser<M>(NodePairLds<NM, 1>{sf + 12 * NM}, NodePairLds<NM, 1>{sf + 14 * NM},
                             NodePairLds<NM>{sf + 16 * NM}, NodePairLds<NM>{sf + 22 * NM}, w0, w1, pc, ps);
        gam_corr = 0.5 * a.qbm * a.qbm * (la.apr * la.apr + la.api * la.api);
    }
    // sqrt(1 + v^2) (+ gam_corr for the laser pushers)
    auto gamma_of = [&](const double *v) {
        double g2 = 1.0 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if constexpr (PGC) g2 = g2 + gam_corr;
        return sqrt(g2);
    };
    double wp0 = ep[0] - bp[1], wp1 = ep[1] + bp[0], wp2 = ep[2];
    double u0[3] = {pp1 * pc + pp2 * ps, pp2 * pc - pp1 * ps, pp3};
    double gam = gamma_of(u0);
    if constexpr (PGC) {
        const double tmp = 0.5 * a.qbm / gam;
        ep[0] = ep[0] - tmp * (la.apr * la.aprg[0] + la.api * la.apig[0]);
        ep[1] = ep[1] - tmp * (la.apr * la.aprg[1] + la.api * la.apig[1]);
        ep[2] = ep[2] + tmp * (la.apr * la.aprg[2] + la.api * la.apig[2]);
    }
    double ut[3], u[3];
    double qtmh2;
    if constexpr (STD) {
        qtmh2 = a.qtmh / (1.0 - a.qbm * psip);
        const double qtmh1 = qtmh2 * gam;
#pragma unroll
        for (int cc = 0; cc < 3; cc++) ut[cc] = u0[cc] + ep[cc] * qtmh1;
    } else if constexpr (PGC) {
        const double qtmh1 = a.qtmh * gam / (gam - u0[2]);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) ut[cc] = u0[cc] + ep[cc] * qtmh1;
        gam = gamma_of(ut);
        qtmh2 = a.qtmh / (gam - ut[2]);
    } else {
        const double qtmh1 = a.qtmh * gam / (gam - u0[2]);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) { ep[cc] = ep[cc] * qtmh1; ut[cc] = u0[cc] + ep[cc]; }
        gam = sqrt(1.0 + ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
        qtmh2 = a.qtmh / (gam - ut[2]);
    }
#pragma unroll
    for (int cc = 0; cc < 3; cc++) bp[cc] = bp[cc] * qtmh2;
    u[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
    u[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
    u[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
    double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) bp[cc] = bp[cc] * ostq;
    ut[0] = ut[0] + u[1] * bp[2] - u[2] * bp[1];
    ut[1] = ut[1] + u[2] * bp[0] - u[0] * bp[2];
    ut[2] = ut[2] + u[0] * bp[1] - u[1] * bp[0];
    if constexpr (STD) {
        // second half kick normalised by the gamma after the rotation
        const double qtmh1 = qtmh2 * gamma_of(ut);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) u[cc] = ut[cc] + ep[cc] * qtmh1;
    } else if constexpr (PGC) {
        const double g1 = gamma_of(ut);
        const double qtmh1 = a.qtmh * g1 / (g1 - ut[2]);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) u[cc] = ut[cc] + ep[cc] * qtmh1;
    } else {
#pragma unroll
        for (int cc = 0; cc < 3; cc++) u[cc] = ut[cc] + ep[cc];
    }
    double du0 = a.idt * (u[0] - u0[0]), du1 = a.idt * (u[1] - u0[1]);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) u[cc] = 0.5 * (u[cc] + u0[cc]);
    double g2 = gamma_of(u);
    double ipsi;
    if constexpr (STD) {
        ipsi = 1.0 / (1.0 - a.qbm * psip);
        if (valid) a.gamma[p] = g2;
    } else {
        ipsi = 1.0 / (g2 - u[2]);
        if (a.store_diag && valid) {
            a.gamma[p] = g2;
            a.psi[p] = (1.0 - 1.0 / ipsi) / a.qbm;
        }
    }
    double dpsi = a.qbm * (wp2 - (wp0 * u[0] + wp1 * u[1]) * ipsi);
    du0 = du0 + u[0] * dpsi * ipsi;
    du1 = du1 + u[1] * dpsi * ipsi;
    t.val[0] = u[0];
    t.val[1] = u[1];
    t.val[2] = u[2];
    t.val[3] = du0;
    t.val[4] = du1;
    t.val[5] = u[0] * u[0] * ipsi;
    t.val[6] = u[0] * u[1] * ipsi;
    t.val[7] = u[1] * u[1] * ipsi;
    double phr = 1.0 * (valid ? qp : 0.0) * ipsi, phi = 0.0;
    t.w[0] = w0;
    t.w[1] = w1;
    t.ph[0] = phr;
    const double ph0r = pc, ph0i = -ps;
#pragma unroll
    for (int m = 1; m <= M; m++) {
        double tr = phr * ph0r - phi * ph0i, ti = phr * ph0i + phi * ph0r;
        phr = tr;
        phi = ti;
        t.ph[2 * m - 1] = phr;
        t.ph[2 * m] = phi;
    }
}

// The matrix form of a wave's cell deposit: the cell's values are the
// 16 x NM product D[n * 8 + comp][k] = sum_p (w[n] val[comp])_p ph[k]_p, a
// GEMM with the particles as the inner dimension, which
// v_mfma_f64_16x16x4f64 accumulates four particles at a time (FP64 on the
// matrix cores; no cross-lane reduction and four accumulators per lane
// instead of 16 NM).  Operand layout of the instruction: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[4 r + l / 16][l % 16]
// in its accumulator r.  The pass's particle terms reach that layout
// through the wave's own LDS rows ([particle][A 0..15, ph 0..NM-1, 0],
// stride mf_stride; lanes past the NM columns read the zero), one half-wave
// of 32 particles at a time; each half's operands are read before its eight
// instructions are issued.
typedef double dep_acc4 __attribute__((ext_vector_type(4)));

template <int NM>
constexpr int mf_stride() { return (17 + NM) | 1; }

// LDS writes of this wave visible to its other lanes (waves of a workgroup
// run different pass counts, so no workgroup barrier here)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one accumulator chain per wave: independent chains (summed at the end)
// would overlap the instruction's latency, but their registers cost a wave
// per SIMD, which measured slower on MI355X (C2: 25.7 against 27.5 us with
// four chains, C3: 23.7 against 18.6 us)
template <int M>
__device__ __forceinline__ void mf_accumulate(double *rows, const PartTerms<M> &t, int lane, dep_acc4 &acc)
{
    constexpr int NM = 2 * M + 1, STR = mf_stride<NM>();
    const int gi = lane >> 4, i = lane & 15;
    const int fb = 16 + min(i, NM);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if ((lane >> 5) == h) {
            double *mine = rows + (lane & 31) * STR;
#pragma unroll
            for (int n = 0; n < 2; n++)
#pragma unroll
                for (int c = 0; c < 8; c++) mine[n * 8 + c] = t.w[n] * t.val[c];
#pragma unroll
            for (int k = 0; k < NM; k++) mine[16 + k] = t.ph[k];
            mine[16 + NM] = 0.0;
        }
        wave_lds_sync();
        double av[8], bv[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const double *r = rows + (4 * s + gi) * STR;
            av[s] = r[i];
            bv[s] = r[fb];
        }
#pragma unroll
        for (int s = 0; s < 8; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
        // all sixteen reads ahead of the eight instructions
        __builtin_amdgcn_sched_group_barrier(0x0100, 16, 0);
        __builtin_amdgcn_sched_group_barrier(0x0008, 8, 0);
        wave_lds_sync();
    }
}

// Halving steps of the wave reduce-scatter applied to each pass's
// contributions before they are accumulated: a lane keeps 16NM >> S
// accumulators instead of 16NM, and the remaining 4 - S steps run once per
// chunk.  Each step costs a chain of dependent cross-lane moves per pass, so
// it pays only where the full set of accumulators would leave one wave per
// SIMD: measured on MI355X (C2, M = 2: 20.6 us per launch with S = 0 and two
// waves per SIMD, 23.1 us with S = 3 and four; C3, M = 4: 41.0 us with
// S = 0 and one wave, 28.0 us with S = 1 and two).
template <int M>
constexpr int dep_pass_steps()
{
    return M <= 2 ? 0 : 1;
}

// the arguments of every species of one launch; workgroup row blockIdx.y
// deposits species blockIdx.y
struct JDepMulti {
    JDepArgs s[kMaxSpecies];
};

// SPEC: a speculative launch of iteration >= 2 of a slice, which returns at
// once when the iteration has already converged (a separate instance, so
// that profiles show the first-iteration launches, which always work, apart)
// MF: the matrix form of the accumulation (mf_accumulate) instead of the
// reduce-scatter; the same values up to the order of the sums
template <int M, bool SPEC, int CPW, bool STD, bool PGC = false, bool MF = false>
__global__ __launch_bounds__(64 * CPW) void k_amjdeposit(const JDepMulti m)
{
    const JDepArgs a = m.s[__builtin_amdgcn_readfirstlane(blockIdx.y)];
    constexpr int NM = 2 * M + 1;
    constexpr int NV = 8 * NM;
    constexpr int NF = dep_staged<NM, PGC>();   // staged field values per cell
    // the fields at the two nodes of each wave's cell (dep_staged layout)
    __shared__ double sf[CPW][NF];
    // the matrix form's particle rows, per wave
    __shared__ double mrows[MF ? CPW * 32 * mf_stride<NM>() : 1];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (dep_group_idle<CPW>(a.nr, a.os)) return;
    const DepChunk ch = dep_chunk<CPW>(a.nr, a.cs, a.os, a.desc, a.n_ov_max, wave);
    const int c = ch.cell;
    const int nm = a.nm, ld = a.ld;
    const int ix = c + 1;
    // every independent load of the prologue first (one round trip): the
    // convergence state, the cell's particle range and its field values
    int done = 0;
    double co = 0.0, cd = 0.0;
    int iters_now = 0;
    if (SPEC) {
        done = a.it->done;
        iters_now = a.it->iters;
        conv_load(a.conv_prev, a.nblk, co, cd);
    }
    const int p0 = ch.pb, p1 = ch.pe;
    double fv[(NF + 63) / 64];
#pragma unroll
    for (int u = 0; u < (NF + 63) / 64; u++) {
        const int l = min(lane + 64 * u, NF - 1);
        if constexpr (PGC) {
            fv[u] = dep_stage_load_pgc<NM>(a.e, a.b, a.ar, a.ai, a.arg, a.aig, nm, ld, l, ix);
        } else {
            const int fb = l / (6 * NM), r = l % (6 * NM);
            const int n = r / (3 * NM), comp = (r / NM) % 3, k = r % NM;
            const double *f = fb ? a.b : a.e;
            fv[u] = f[((size_t)comp * nm + k) * ld + ix + n];
        }
    }
    if (SPEC) {
        int stop = done;
        if (!stop) {
            // 'compare' of the previous iteration (simulation_class.f03:396):
            // every wave evaluates it (the same decision everywhere);
            // workgroup 0 records it
            double rel, ab;
            conv_finish(co, cd, rel, ab);
            stop = rel < a.reltol || ab < a.abstol;
            if (stop && blockIdx.x == 0 && threadIdx.x == 0) a.it->done = 1;
        }
        if (a.report && blockIdx.x == 0 && threadIdx.x == 0) {
            a.report->done = stop;
            a.report->iters = iters_now;
        }
        if (stop) return;
    }
#pragma unroll
    for (int u = 0; u < (NF + 63) / 64; u++) {
        const int l = lane + 64 * u;
        if (l < NF) sf[wave][l] = fv[u];
    }
    __syncthreads();
    if (!ch.live) return;
    const double *__restrict__ X1 = a.x1;
    const double *__restrict__ X2 = a.x2;
    const double *__restrict__ PP1 = a.p1;
    const double *__restrict__ PP2 = a.p2;
    const double *__restrict__ PP3 = a.p3;
    const double *__restrict__ Q = a.q;
    constexpr int S = dep_pass_steps<M>();
    constexpr int NA = MF ? 1 : (2 * NV) >> S;
    double acc[NA];
#pragma unroll
    for (int v = 0; v < NA; v++) acc[v] = 0.0;
    dep_acc4 macc = {0.0, 0.0, 0.0, 0.0};
    double *const wrows = mrows + (MF ? wave * 32 * mf_stride<NM>() : 0);
    // every pass has all 64 lanes active (the exchanges need them); lanes
    // past the chunk's end deposit zero.  Software-pipelined: the next pass's
    // particle loads are in flight while this pass computes.
    const int npass = __builtin_amdgcn_readfirstlane((p1 - p0 + 63) >> 6);
    const int pl = max(p1 - 1, p0);
    int p = p0 + lane;
    int pc = min(p, pl);
    double x1 = X1[pc], x2 = X2[pc], pp1 = PP1[pc], pp2 = PP2[pc], pp3 = PP3[pc], qp = Q[pc];
    double psip = STD ? a.psi[pc] : 0.0;
    for (int k = 0; k < npass; k++, p += 64) {
        // the staged fields are re-read from LDS (broadcast) every particle
        // rather than held in registers: the accumulators need the VGPRs
        __asm__ volatile("" ::: "memory");
        const int pn = min(p + 64, pl);
        const double nx1 = X1[pn], nx2 = X2[pn], np1 = PP1[pn], np2 = PP2[pn], np3 = PP3[pn], nq = Q[pn];
        const double npsi = STD ? a.psi[pn] : 0.0;
        PartTerms<M> pt;
        particle_terms<M, STD, PGC>(a, sf[wave], c, p, p < p1, x1, x2, pp1, pp2, pp3, qp, psip, pt);
        if constexpr (MF) {
            mf_accumulate<M>(wrows, pt, lane, macc);
        } else {
            const DepTerms<M> t(pt);
#pragma unroll
            for (int i = 0; i < NA; i++) acc[i] += reduced_value<NM, S>(t, i, lane);
        }
        x1 = nx1; x2 = nx2; pp1 = np1; pp2 = np2; pp3 = np3; qp = nq; psip = npsi;
    }
    double *D0 = ch.ovf ? a.O0 : a.P0, *D1 = ch.ovf ? a.O1 : a.P1;
    if constexpr (MF) {
        // lane l holds D[4 r + l / 16][l % 16]: row = node * 8 + comp, column = mode array
        const int k = lane & 15, rb = lane >> 4;
        const dep_acc4 sum = macc;
        if (k < NM) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 4 * r + rb, comp = row & 7;
                double *D = row < 8 ? D0 : D1;
                D[(size_t)ch.row * NV + comp * NM + k] = sum[r];
            }
        }
    } else {
        wave_reduce_from<NM, S>(acc);
        const int g = lane >> 2;
#pragma unroll
        for (int i = 0; i < NM; i++) {
            const int v = g * NM + i;
            if ((lane & 3) == (i & 3)) {
                if (v < NV) D0[(size_t)ch.row * NV + v] = acc[i];
                else D1[(size_t)ch.row * NV + v - NV] = acc[i];
            }
        }
    }
}

void time_begin(qpad_ctx *c, int which, hipEvent_t *stop);
void time_end(qpad_ctx *c, int which, hipEvent_t stop);

static JDepArgs jdep_args(qpad_ctx *c, int s, int l, bool store_diag, bool report)
{
    const bool speculative = l >= 2;
    Species &sp = c->sp[s];
    JDepArgs a;
    a.nr = c->g.nr;
    a.ld = c->g.ld;
    a.nm = c->g.nm;
    a.idr = 1.0 / c->g.dr;
    a.qbm = sp.qbm;
    a.qtmh = 0.5 * sp.qbm * c->dt2d;
    a.idt = 1.0 / c->dt2d;
    a.x1 = sp.x1; a.x2 = sp.x2; a.p1 = sp.p1; a.p2 = sp.p2; a.p3 = sp.p3; a.q = sp.q;
    a.gamma = sp.gamma;
    a.psi = sp.psi;
    a.store_diag = store_diag;
    a.cs = sp.cell_start;
    a.e = c->f[QPAD_F_E];
    a.b = c->f[QPAD_F_B];
    a.ar = c->f[QPAD_F_AR];
    a.ai = c->f[QPAD_F_AI];
    a.arg = c->f[QPAD_F_AR_GRAD];
    a.aig = c->f[QPAD_F_AI_GRAD];
    a.P0 = sp.P0;
    a.P1 = sp.P1;
    a.it = c->it;
    a.nblk = conv_blocks(c->g.nr);
    a.conv_prev = c->conv_blk + (size_t)((l - 1) & 1) * 2 * a.nblk;
    a.reltol = c->g.reltol;
    a.abstol = c->g.abstol;
    a.report = (report && speculative) ? c->it_report : nullptr;
    a.os = sp.ov_start;
    a.desc = sp.ov_desc;
    a.O0 = sp.O0;
    a.O1 = sp.O1;
    a.n_ov_max = sp.n_ov_max;
    return a;
}

// the species of species-list entries in one launch (grid rows by list
// entry), all of the push type STD, PGC names
template <bool STD, bool PGC>
static int launch_jdep_list(qpad_ctx *c, const int *list, int n, int l, bool store_diag, bool timed, bool report)
{
    const bool speculative = l >= 2;
    JDepMulti m;
    memset(&m, 0, sizeof(m));
    const int cpw = c->dep_cells;
    int gx = 1;
    for (int i = 0; i < n; i++) {
        const int s = list[i];
        // the report (the host's copy of the decision) from the first entry only
        m.s[i] = jdep_args(c, s, l, store_diag, report && i == 0);
        gx = std::max(gx, dep_grid(c->g.nr, c->sp[s].npp, c->sp[s].n_ov_max, cpw));
    }
    dim3 grid((unsigned)gx, (unsigned)n), blk(64 * cpw);
    hipEvent_t stop = nullptr;
    if (timed) time_begin(c, 0, &stop);
    const bool mf = c->dep_mfma != 0;
#define QPAD_DEP3(MM, CC, MFV)                                                                                \
    do {                                                                                                      \
        if (speculative)                                                                                      \
            hipLaunchKernelGGL((k_amjdeposit<MM, true, CC, STD, PGC, MFV>), grid, blk, 0, c->stream, m);       \
        else hipLaunchKernelGGL((k_amjdeposit<MM, false, CC, STD, PGC, MFV>), grid, blk, 0, c->stream, m);     \
    } while (0)
#define QPAD_DEP2(MM, CC)                    \
    do {                                     \
        if (mf) QPAD_DEP3(MM, CC, true);     \
        else QPAD_DEP3(MM, CC, false);       \
    } while (0)
#define QPAD_DEP(MM)                                 \
    do {                                             \
        if (cpw == 2) QPAD_DEP2(MM, 2);              \
        else if (cpw == 8) QPAD_DEP2(MM, 8);         \
        else QPAD_DEP2(MM, 4);                       \
    } while (0)
    switch (c->g.M) {
    case 0: QPAD_DEP(0); break;
    case 1: QPAD_DEP(1); break;
    case 2: QPAD_DEP(2); break;
    case 3: QPAD_DEP(3); break;
    default: QPAD_DEP(4); break;
    }
#undef QPAD_DEP
#undef QPAD_DEP2
#undef QPAD_DEP3
    QPAD_HIP_CHECK(hipGetLastError());
    if (timed) time_end(c, 0, stop);
    return 0;
}

// the four pushers in launch order: robust, standard, robust_pgc,
// standard_pgc
static int push_group(int push_type)
{
    switch (push_type) {
    case QPAD_PUSH2_STD: return 1;
    case QPAD_PUSH2_ROBUST_PGC: return 2;
    case QPAD_PUSH2_STD_PGC: return 3;
    default: return 0;
    }
}

// species s0 .. s1-1: one launch per push type present (the first launch
// carries the timing and the host report when there is one)
static int launch_jdep(qpad_ctx *c, int s0, int s1, int l, bool store_diag, bool timed, bool report)
{
    int list[4][kMaxSpecies], n[4] = {0, 0, 0, 0};
    for (int s = s0; s < s1; s++) {
        const int g = push_group(c->sp[s].cfg.push_type);
        list[g][n[g]++] = s;
    }
    bool first = true;
    int rc = 0;
    for (int g = 0; g < 4 && !rc; g++) {
        if (!n[g]) continue;
        const bool tm = timed && first, rp = report && first;
        if (g == 0) rc = launch_jdep_list<false, false>(c, list[g], n[g], l, store_diag, tm, rp);
        else if (g == 1) rc = launch_jdep_list<true, false>(c, list[g], n[g], l, store_diag, tm, rp);
        else if (g == 2) rc = launch_jdep_list<false, true>(c, list[g], n[g], l, store_diag, tm, rp);
        else rc = launch_jdep_list<true, true>(c, list[g], n[g], l, store_diag, tm, rp);
        first = false;
    }
    return rc;
}

int launch_amjdeposit(qpad_ctx *c, int s, int l, bool store_diag, bool timed, bool report)
{
    return launch_jdep(c, s, s + 1, l, store_diag, timed, report);
}

int launch_amjdeposit_all(qpad_ctx *c, int l, bool store_diag, bool timed, bool report)
{
    if (c->nspecies == 0) return 0;
    return launch_jdep(c, 0, c->nspecies, l, store_diag, timed, report);
}

// ---------------------------------------------------------------------------
// push_u_robust (part2d_class.f03:1879-1965) + push_x (:2221-2262) + the new
// sort key (update_bound :2307-2353 is the key nr).  Thread per particle;
// particles that already left the box (past cell_start[nr]) keep key nr.
// ---------------------------------------------------------------------------
struct PushArgs {
    int nr, ld, nm;
    int64_t nslots;
    double idr, edge, qtmh, dt, qbm;
    double clamp;        // expansion-factor clamp of a sub-cycled push (0: none)
    double *x1, *x2, *p1, *p2, *p3, *gamma;
    const double *psi;   // interpolated psi (standard pusher)
    const int32_t *cs;
    uint32_t *key;
    const double *e, *b;
    const double *ar, *ai, *arg, *aig;   // laser envelope and gradients (PGC)
};

// the arguments of every species of one launch (workgroup rows by species)
struct PushMulti {
    PushArgs s[kMaxSpecies];
};

// STD: push_u_std_part2d (part2d_class.f03:1790-1877), normalised by the
// interpolated psi and the gamma the last current deposit stored, where
// push_u_robust_part2d (:1879-1965) uses the particle's own gamma - p_z.
// PGC: push_u_robust_pgc_part2d / push_u_std_pgc_part2d (:1967-2219, the
// same steps): the standard normalisation, the ponderomotive term in E with
// the stored gamma, and the new gamma with gam_corr extrapolated half a step
// in xi (the envelope's xi gradient)
template <int M, bool STD, bool PGC = false>
__global__ __launch_bounds__(256) void k_push(const PushMulti m)
{
    const PushArgs a = m.s[__builtin_amdgcn_readfirstlane(blockIdx.y)];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nslots) return;
    const int npp = a.cs[a.nr];
    if (p >= npp) return;
    const int nm = a.nm, ld = a.ld;
    double x1 = a.x1[p], x2 = a.x2[p];
    double r = radius(x1, x2);
    const double pc = x1 / r, ps = x2 / r;
    double pos = __dmul_rn(r, a.idr);
    if (!(pos >= 0.0)) pos = 0.0;   // r = 0 or NaN: keep the gather in bounds
    int ic = (int)fmin(pos, (double)(a.nr - 1));
    const int ix = ic + 1;
    const double w1 = pos - (double)ic, w0 = 1.0 - w1;
    double ep[3], bp[3];
    gather3<M>(NodePair{a.e, nm, ld, ix}, w0, w1, pc, ps, ep);
    gather3<M>(NodePair{a.b, nm, ld, ix}, w0, w1, pc, ps, bp);
    double t = ep[0] * pc - ep[1] * ps;
    ep[1] = ep[0] * ps + ep[1] * pc;
    ep[0] = t;
    t = bp[0] * pc - bp[1] * ps;
    bp[1] = bp[0] * ps + bp[1] * pc;
    bp[0] = t;
    double gam_corr = 0.0, apr_z = 0.0, api_z = 0.0, apr = 0.0, api = 0.0;
    if constexpr (PGC) {
        LaserAt la = gather_laser<M>(NodePair{a.ar, nm, ld, ix}, NodePair{a.ai, nm, ld, ix},
                                     NodePair{a.arg, nm, ld, ix}, NodePair{a.aig, nm, ld, ix}, w0, w1, pc, ps);
        t = la.aprg[0] * pc - la.aprg[1] * ps;
        la.aprg[1] = la.aprg[0] * ps + la.aprg[1] * pc;
        la.aprg[0] = t;
        t = la.apig[0] * pc - la.apig[1] * ps;
        la.apig[1] = la.apig[0] * ps + la.apig[1] * pc;
        la.apig[0] = t;
        const double qbm2_hf = a.qbm * a.qbm * 0.5;
        gam_corr = qbm2_hf * (la.apr * la.apr + la.api * la.api);
        const double tmp = 0.5 * a.qbm / a.gamma[p];
        ep[0] = ep[0] - tmp * (la.apr * la.aprg[0] + la.api * la.apig[0]);
        ep[1] = ep[1] - tmp * (la.apr * la.aprg[1] + la.api * la.apig[1]);
        ep[2] = ep[2] + tmp * (la.apr * la.aprg[2] + la.api * la.apig[2]);
        apr = la.apr;
        api = la.api;
        apr_z = la.aprg[2];
        api_z = la.apig[2];
    }
    double pp[3] = {a.p1[p], a.p2[p], a.p3[p]}, ut[3];
    double qtmh1, qtmh2;
    if constexpr (STD || PGC) {
        qtmh1 = a.qtmh / (1.0 - a.qbm * a.psi[p]);
        qtmh2 = qtmh1 * a.gamma[p];
    } else {
        const double gam = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
        qtmh1 = a.qtmh / (gam - pp[2]);
        qtmh2 = qtmh1 * gam;
    }
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
        ep[cc] = ep[cc] * qtmh2;
        bp[cc] = bp[cc] * qtmh1;
        ut[cc] = pp[cc] + ep[cc];
    }
    pp[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
    pp[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
    pp[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
    double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) bp[cc] = bp[cc] * ostq;
    ut[0] = ut[0] + pp[1] * bp[2] - pp[2] * bp[1];
    ut[1] = ut[1] + pp[2] * bp[0] - pp[0] * bp[2];
    ut[2] = ut[2] + pp[0] * bp[1] - pp[1] * bp[0];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) pp[cc] = ut[cc] + ep[cc];
    double g2;
    if constexpr (PGC) {
        const double qbm2_hf = a.qbm * a.qbm * 0.5;
        double tmp = apr_z * a.dt;
        gam_corr = gam_corr + qbm2_hf * (apr + 0.25 * tmp) * tmp;
        tmp = api_z * a.dt;
        gam_corr = gam_corr + qbm2_hf * (api + 0.25 * tmp) * tmp;
        g2 = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2] + gam_corr);
    } else {
        g2 = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
    }
    if (a.clamp > 0.0) {
        // clamp_exp_fac (proj_subcyc/part2d_subcyc_class.f03:46-66): the
        // momentum scaled so that gamma / (gamma - p_z) becomes the clamp
        const double F = a.clamp;
        if (g2 / (g2 - pp[2]) > F) {
            double sc = (F - 1.0) * (F - 1.0);
            sc = sc / (F * F * pp[2] * pp[2] - sc * (g2 * g2 - 1.0));
            sc = sqrt(sc);
            pp[0] = pp[0] * sc;
            pp[1] = pp[1] * sc;
            pp[2] = pp[2] * sc;
            g2 = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
        }
    }
    double dtc = a.dt / (g2 - pp[2]);
    x1 = x1 + pp[0] * dtc;
    x2 = x2 + pp[1] * dtc;
    a.p1[p] = pp[0];
    a.p2[p] = pp[1];
    a.p3[p] = pp[2];
    a.gamma[p] = g2;
    a.x1[p] = x1;
    a.x2[p] = x2;
    a.key[p] = cell_key(x1, x2, a.idr, a.edge, a.nr);
}

static PushArgs push_args(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    PushArgs a;
    a.nr = c->g.nr;
    a.ld = c->g.ld;
    a.nm = c->g.nm;
    a.nslots = sp.npp;
    a.idr = 1.0 / c->g.dr;
    a.edge = (double)c->g.nr * c->g.dr;
    a.qtmh = sp.qbm * c->dt2d * 0.5;
    a.dt = c->dt2d;
    a.clamp = c->clamp2d;
    a.qbm = sp.qbm;
    a.x1 = sp.x1; a.x2 = sp.x2; a.p1 = sp.p1; a.p2 = sp.p2; a.p3 = sp.p3; a.gamma = sp.gamma;
    a.psi = sp.psi;
    a.cs = sp.cell_start;
    a.key = sp.key;
    a.e = c->f[QPAD_F_E];
    a.b = c->f[QPAD_F_B];
    a.ar = c->f[QPAD_F_AR];
    a.ai = c->f[QPAD_F_AI];
    a.arg = c->f[QPAD_F_AR_GRAD];
    a.aig = c->f[QPAD_F_AI_GRAD];
    return a;
}

// the species of list entries in one launch (species without particles get
// idle rows), all of the push type STD, PGC names
template <bool STD, bool PGC>
static int launch_push_list(qpad_ctx *c, const int *list, int n)
{
    PushMulti m;
    memset(&m, 0, sizeof(m));
    int64_t nmax = 0;
    for (int i = 0; i < n; i++) {
        m.s[i] = push_args(c, list[i]);
        nmax = std::max(nmax, c->sp[list[i]].npp);
    }
    if (nmax == 0) return 0;
    dim3 grid((unsigned)((nmax + 255) / 256), (unsigned)n), blk(256);
    hipEvent_t stop = nullptr;
    time_begin(c, 1, &stop);
    switch (c->g.M) {
    case 0: hipLaunchKernelGGL((k_push<0, STD, PGC>), grid, blk, 0, c->stream, m); break;
    case 1: hipLaunchKernelGGL((k_push<1, STD, PGC>), grid, blk, 0, c->stream, m); break;
    case 2: hipLaunchKernelGGL((k_push<2, STD, PGC>), grid, blk, 0, c->stream, m); break;
    case 3: hipLaunchKernelGGL((k_push<3, STD, PGC>), grid, blk, 0, c->stream, m); break;
    default: hipLaunchKernelGGL((k_push<4, STD, PGC>), grid, blk, 0, c->stream, m); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());
    time_end(c, 1, stop);
    return 0;
}

// species s0 .. s1-1: one launch per push present (the two PGC pushers
// push alike)
static int launch_push_range(qpad_ctx *c, int s0, int s1)
{
    int list[3][kMaxSpecies], n[3] = {0, 0, 0};
    for (int s = s0; s < s1; s++) {
        const int g = std::min(push_group(c->sp[s].cfg.push_type), 2);
        list[g][n[g]++] = s;
    }
    int rc = 0;
    if (n[0]) rc = launch_push_list<false, false>(c, list[0], n[0]);
    if (!rc && n[1]) rc = launch_push_list<true, false>(c, list[1], n[1]);
    if (!rc && n[2]) rc = launch_push_list<false, true>(c, list[2], n[2]);
    return rc;
}

// ---------------------------------------------------------------------------
// interp_psi_part2d (part2d_class.f03:2264-2305): the field psi at every
// particle of a standard-pusher species, after the slice's psi solve
// (simulation_class.f03:357-359).  Thread per particle slot; particles past
// cell_start[nr] have left the box.
// ---------------------------------------------------------------------------
struct InterpPsiArgs {
    int nr, ld;
    int64_t nslots;
    double idr;
    const double *x1, *x2, *f;
    double *psi;
    const int32_t *cs;
};

template <int M>
__global__ __launch_bounds__(256) void k_interp_psi(const InterpPsiArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nslots || p >= a.cs[a.nr]) return;
    constexpr int NM = 2 * M + 1;
    const double x1 = a.x1[p], x2 = a.x2[p];
    const double r = radius(x1, x2);
    const double pc = x1 / r, ps = x2 / r;
    double pos = __dmul_rn(r, a.idr);
    if (!(pos >= 0.0)) pos = 0.0;
    const int ic = (int)fmin(pos, (double)(a.nr - 1));
    const int ix = ic + 1;
    const double w1 = pos - (double)ic, w0 = 1.0 - w1;
    // the gather of interp_field (interp_part2d.f03:66-110) for one
    // component, in gather3's order of operations
    const double *f = a.f + ix;
    double v = 0.0;
    v = v + f[0] * w0;
    v = v + f[1] * w1;
    double phr = 1.0, phi = 0.0;
#pragma unroll
    for (int m = 1; m <= M; m++) {
        const double tr = phr * pc - phi * ps, ti = phr * ps + phi * pc;
        phr = tr;
        phi = ti;
        const double ph_re = 2.0 * phr, ph_im = 2.0 * phi;
        const double *fr = f + (size_t)(2 * m - 1) * a.ld, *fi = f + (size_t)(2 * m) * a.ld;
        v = v + (fr[0] * ph_re - fi[0] * ph_im) * w0;
        v = v + (fr[1] * ph_re - fi[1] * ph_im) * w1;
    }
    (void)NM;
    a.psi[p] = v;
}

int launch_interp_psi(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    const int pt = sp.cfg.push_type;
    if ((pt != QPAD_PUSH2_STD && pt != QPAD_PUSH2_STD_PGC) || sp.npp == 0) return 0;
    InterpPsiArgs a;
    a.nr = c->g.nr;
    a.ld = c->g.ld;
    a.nslots = sp.npp;
    a.idr = 1.0 / c->g.dr;
    a.x1 = sp.x1;
    a.x2 = sp.x2;
    a.f = c->f[QPAD_F_PSI];
    a.psi = sp.psi;
    a.cs = sp.cell_start;
    const dim3 grid((unsigned)((sp.npp + 255) / 256)), blk(256);
    switch (c->g.M) {
    case 0: hipLaunchKernelGGL(k_interp_psi<0>, grid, blk, 0, c->stream, a); break;
    case 1: hipLaunchKernelGGL(k_interp_psi<1>, grid, blk, 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL(k_interp_psi<2>, grid, blk, 0, c->stream, a); break;
    case 3: hipLaunchKernelGGL(k_interp_psi<3>, grid, blk, 0, c->stream, a); break;
    default: hipLaunchKernelGGL(k_interp_psi<4>, grid, blk, 0, c->stream, a); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_push(qpad_ctx *c, int s) { return launch_push_range(c, s, s + 1); }

int launch_push_all(qpad_ctx *c) { return c->nspecies ? launch_push_range(c, 0, c->nspecies) : 0; }

// q_spe += qn (the neutralising background of species s)
__global__ void k_axpy(double *y, const double *x, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = y[t] + x[t];
}

// ---------------------------------------------------------------------------
// species2d%q and cbq (species2d_class.f03:186-208, :396-410): the species'
// own rho - Jz (its deposit with the axis rules and 1/r; the neutralising
// background enters only the total), then J_z of its last executed current
// deposit added, and the sum copied into the species' history slice
// ---------------------------------------------------------------------------
int launch_species_q_start(qpad_ctx *c, int s)
{
    return launch_combine_q(c, s, c->sp[s].qsl, 1.0, false);
}

// thread per (node i = 0..nr+1, mode array k): q += J_z of the species
// (comp 2 of its current-deposit partials, axis x8 at m = 0 and 0 above),
// then history slice = q
__global__ void k_species_cbq(const double *P0, const double *P1, const double *O0, const double *O1,
                              const int32_t *os, double *q, double *hist, int nr, int nm, int ld)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nm * (nr + 2)) return;
    const int i = (int)(tid % (nr + 2)), k = (int)(tid / (nr + 2));
    const int NV = 8 * nm, v = 2 * nm + k;
    double val = 0.0;
    if (i >= 1 && i <= nr) val += cell_partial(P0, O0, os, i - 1, NV, v);
    if (i >= 2) val += cell_partial(P1, O1, os, i - 2, NV, v);
    if (i == 0) val = 0.0;
    else if (i == 1) val = (k == 0) ? 8.0 * val : 0.0;
    else val = val * (1.0 / (double)(i - 1));
    const size_t o = (size_t)k * ld + i;
    const double x = q[o] + val;
    q[o] = x;
    hist[o] = x;
}

int launch_species_cbq(qpad_ctx *c, int s, int j)
{
    const Species &sp = c->sp[s];
    const int64_t n = (int64_t)c->g.nm * (c->g.nr + 2);
    double *hist = sp.h_q + (size_t)j * c->g.nm * c->g.ld;
    hipLaunchKernelGGL(k_species_cbq, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, sp.P0, sp.P1,
                       sp.O0, sp.O1, sp.ov_start, sp.qsl, hist, c->g.nr, c->g.nm, c->g.ld);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_add_qn_into_qspe(qpad_ctx *c, int s)
{
    int64_t n = (int64_t)c->g.nm * c->g.ld;
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->f[QPAD_F_Q_SPE],
                       c->sp[s].qn, n);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace qpad
