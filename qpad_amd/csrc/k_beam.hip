// k_beam.hip -- 3D beam on the GPU (FP64): random Gaussian injection,
// (r, xi) charge deposit into the q_beam history and the reduced pusher.
//
// Beam particles are sorted by their (xi, r) cell before the deposit; every
// history node then sums the particles of its four neighbouring cells in a
// fixed order (node-centric gather, deterministic, no atomics).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "kernels.h"
#include "sort_config.h"
#include "rng.h"

namespace qpad {

__device__ inline double bradius(double x1, double x2)
{
    return sqrt(__dadd_rn(__dmul_rn(x1, x1), __dmul_rn(x2, x2)));
}

// ---------------------------------------------------------------------------
// injection: fdist3d_rnd_class.f03:360-513 (Cartesian Gaussian profile)
// ---------------------------------------------------------------------------
struct BeamGenArgs {
    int64_t n;
    uint64_t seed;
    double center[3], sigma[3], lo[3], hi[3], uth[3];
    double gamma, edge_r, edge_z, qpp;
    double *x1, *x2, *x3, *p1, *p2, *p3, *q;
    uint32_t *flag;
};

__global__ void k_beam_gen(BeamGenArgs a)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double xt[3] = {0.0, 0.0, 0.0};
    // redraw until inside range1..range3 (bounded: the probability of 4096
    // consecutive rejections is nil for ranges of a few sigma)
    for (uint32_t att = 0; att < 4096u; att++) {
        double g[4];
        normal4(a.seed, (uint64_t)i, 2u * att, g);
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            xt[d] = g[d] * a.sigma[d] + a.center[d];
            if (xt[d] < a.lo[d] || xt[d] >= a.hi[d]) ok = false;
        }
        if (ok) break;
    }
    double r = sqrt(xt[0] * xt[0] + xt[1] * xt[1]);
    bool keep = (r < a.edge_r) && (xt[2] >= 0.0) && (xt[2] < a.edge_z);
    double g[4];
    normal4(a.seed, (uint64_t)i, 0x80000000u, g);
    double p1 = a.uth[0] * g[0], p2 = a.uth[1] * g[1];
    double p3 = a.uth[2] * g[2] + a.gamma;
    p3 = sqrt(p3 * p3 - p1 * p1 - p2 * p2 - 1.0);
    a.x1[i] = xt[0];
    a.x2[i] = xt[1];
    a.x3[i] = xt[2];
    a.p1[i] = p1;
    a.p2[i] = p2;
    a.p3[i] = p3;
    a.q[i] = a.qpp;
    a.flag[i] = keep ? 1u : 0u;
}

struct CompactArgs {
    int64_t n;
    const uint32_t *flag, *pos;
    const double *s[7];
    double *d[7];
};

__global__ void k_beam_compact(CompactArgs a)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n || !a.flag[i]) return;
    uint32_t o = a.pos[i];
    for (int f = 0; f < 7; f++) a.d[f][o] = a.s[f][i];
}

// quiet start: mirrored copies (-x, -y, -p_perp), charge halved
__global__ void k_beam_quiet(double *x1, double *x2, double *x3, double *p1, double *p2, double *p3,
                             double *q, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x1[n + i] = -x1[i];
    x2[n + i] = -x2[i];
    x3[n + i] = x3[i];
    p1[n + i] = -p1[i];
    p2[n + i] = -p2[i];
    p3[n + i] = p3[i];
    double qq = 0.5 * q[i];
    q[i] = qq;
    q[n + i] = qq;
}

int launch_beam_inject(qpad_ctx *c, int ib)
{
    Beam &bm = c->bm[ib];
    const qpad_beam_cfg &cf = bm.cfg;
    const double pi = 4.0 * atan(1.0);
    int64_t n = cf.total_num;
    if (n <= 0) {
        bm.npp = 0;
        return 0;
    }
    BeamGenArgs a;
    a.n = n;
    a.seed = cf.seed;
    for (int d = 0; d < 3; d++) {
        a.center[d] = cf.center[d];
        a.sigma[d] = cf.sigma[d];
        a.lo[d] = cf.range_lo[d];
        a.hi[d] = cf.range_hi[d];
        a.uth[d] = cf.uth[d];
    }
    a.gamma = cf.gamma;
    a.edge_r = (double)c->g.nr * c->g.dr;
    a.edge_z = (double)c->g.nxi * c->g.dxi;
    a.qpp = copysign(1.0, cf.qm) * cf.total_charge /
            (2.0 * pi * (double)n * c->g.dr * c->g.dr * c->g.dxi);
    a.x1 = bm.a_x1; a.x2 = bm.a_x2; a.x3 = bm.a_x3;
    a.p1 = bm.a_p1; a.p2 = bm.a_p2; a.p3 = bm.a_p3; a.q = bm.a_q;
    a.flag = bm.key_alt;
    unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_beam_gen, dim3(nb), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    size_t bytes = bm.sort_tmp_bytes;
    QPAD_HIP_CHECK(rocprim::exclusive_scan(bm.sort_tmp, bytes, bm.key_alt, bm.idx_alt, 0u, (size_t)n,
                                           rocprim::plus<uint32_t>(), c->stream));
    CompactArgs ca;
    ca.n = n;
    ca.flag = bm.key_alt;
    ca.pos = bm.idx_alt;
    const double *src[7] = {bm.a_x1, bm.a_x2, bm.a_x3, bm.a_p1, bm.a_p2, bm.a_p3, bm.a_q};
    double *dst[7] = {bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q};
    for (int f = 0; f < 7; f++) {
        ca.s[f] = src[f];
        ca.d[f] = dst[f];
    }
    hipLaunchKernelGGL(k_beam_compact, dim3(nb), dim3(256), 0, c->stream, ca);
    QPAD_HIP_CHECK(hipGetLastError());
    uint32_t last_pos = 0, last_flag = 0;
    QPAD_HIP_CHECK(hipMemcpyAsync(&last_pos, bm.idx_alt + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    QPAD_HIP_CHECK(hipMemcpyAsync(&last_flag, bm.key_alt + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    int64_t nk = (int64_t)last_pos + last_flag;
    if (cf.quiet) {
        if (nk > 0) {
            hipLaunchKernelGGL(k_beam_quiet, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, c->stream,
                               bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q, nk);
            QPAD_HIP_CHECK(hipGetLastError());
        }
        nk *= 2;
    }
    bm.npp = nk;
    return 0;
}

// ---------------------------------------------------------------------------
// lattice injection: fdist3d_std_class.f03:439-613.  QPAD walks the lattice
// k (xi cell) > j (theta sector) > i (r cell) > i3 > i2 > i1 and keeps the
// points inside the ranges whose density reaches den_min; the kept points
// are numbered in that order.  Here the walk is restricted to the cells the
// ranges allow (a box of the same order), one thread per lattice point: a
// counting pass flags the kept points, an exclusive scan numbers them, and
// the writing pass stores each one at its number.
// ---------------------------------------------------------------------------
struct LatticeArgs {
    qpad_beam_std_cfg cf;
    int64_t n;              // lattice points in the box
    int k0, ni;             // xi cells k0 .. (1-based), r cells 1 .. ni
    const double *den_tab;  // the density of every box point, profiles aside (null: the profiles)
    double dr, dz, dtheta, coef;
    uint32_t *flag;
    const uint32_t *pos;
    double *x1, *x2, *x3, *p1, *p2, *p3, *q;
};

// density profile of direction d at coordinate x (fdist3d_std_lib.f03)
__device__ double lattice_profile(const qpad_beam_std_cfg &c, int d, double x)
{
    if (c.profile[d] == QPAD_PROF_GAUSSIAN) {
        const double t = (x - c.gauss_center[d]) / c.gauss_sigma[d];
        return exp(-0.5 * (t * t));
    }
    if (c.profile[d] == QPAD_PROF_PIECEWISE) {
        const int n = c.pw_n[d];
        const double *xs = c.pw_x[d], *fx = c.pw_fx[d];
        if (x < xs[0] || x > xs[n - 1]) return 0.0;
        for (int i = 1; i < n; i++)
            if (x <= xs[i]) return fx[i - 1] + (fx[i] - fx[i - 1]) / (xs[i] - xs[i - 1]) * (x - xs[i - 1]);
        return 0.0;
    }
    return 1.0;
}

// lattice point L of the box: position (geometry coordinates), rn, density;
// false when QPAD's loop skips it
__device__ bool lattice_point(const LatticeArgs &a, int64_t L, double xt[3], double *rn, double *den)
{
    const qpad_beam_std_cfg &c = a.cf;
    const int64_t L0 = L;
    const int i1 = (int)(L % c.ppc[0]) + 1;
    L /= c.ppc[0];
    const int i2 = (int)(L % c.ppc[1]) + 1;
    L /= c.ppc[1];
    const int i3 = (int)(L % c.ppc[2]) + 1;
    L /= c.ppc[2];
    const int i = (int)(L % a.ni) + 1;
    L /= a.ni;
    const int j = (int)(L % c.num_theta) + 1;
    const int k = (int)(L / c.num_theta) + a.k0;
    const double zn = ((double)i3 - 0.5) / (double)c.ppc[2] + (double)(k - 1);
    const double theta = (((double)i2 - 0.5) / (double)c.ppc[1] + (double)j - 1.0) * a.dtheta;
    *rn = ((double)i1 - 0.5) / (double)c.ppc[0] + (double)(i - 1);
    if (c.geometry == QPAD_GEOM_CYLINDRICAL) {
        xt[0] = *rn * a.dr;
        xt[1] = theta;
    } else {
        xt[0] = *rn * a.dr * cos(theta);
        xt[1] = *rn * a.dr * sin(theta);
    }
    xt[2] = zn * a.dz;
    for (int d = 0; d < 3; d++)
        if (xt[d] < c.range_lo[d] || xt[d] > c.range_hi[d]) return false;
    if (a.den_tab) {
        *den = a.den_tab[L0] * c.density;
    } else {
        double dv[3];
        for (int d = 0; d < 3; d++) dv[d] = lattice_profile(c, d, xt[d]);
        *den = dv[0] * dv[1] * dv[2] * c.density;
    }
    return !(*den < c.den_min);
}

__global__ void k_lattice_flag(LatticeArgs a)
{
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= a.n) return;
    double xt[3], rn, den;
    a.flag[L] = lattice_point(a, L, xt, &rn, &den) ? 1u : 0u;
}

__global__ void k_lattice_write(LatticeArgs a)
{
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= a.n || !a.flag[L]) return;
    double xt[3], rn, den;
    lattice_point(a, L, xt, &rn, &den);
    const uint32_t ip = a.pos[L];
    const qpad_beam_std_cfg &c = a.cf;
    if (c.geometry == QPAD_GEOM_CYLINDRICAL) {
        a.x1[ip] = xt[0] * cos(xt[1]);
        a.x2[ip] = xt[0] * sin(xt[1]);
    } else {
        a.x1[ip] = xt[0];
        a.x2[ip] = xt[1];
    }
    a.x3[ip] = xt[2];
    double g[4];
    normal4(c.seed, (uint64_t)ip, 0x80000000u, g);
    const double p1 = c.uth[0] * g[0], p2 = c.uth[1] * g[1];
    double p3 = c.uth[2] * g[2] + c.gamma;
    p3 = sqrt(p3 * p3 - p1 * p1 - p2 * p2 - 1.0);
    a.p1[ip] = p1;
    a.p2[ip] = p2;
    a.p3[ip] = p3;
    a.q[ip] = rn * den * a.coef;
}

// Twiss tilt of the transverse phase space, then the xi-dependent offset
// (fdist3d_std_class.f03:580-608), over every particle including the
// quiet-start mirrors
__global__ void k_lattice_shape(qpad_beam_std_cfg c, int64_t n, double *x1, double *x2, const double *x3,
                                double *p1, double *p2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (c.alpha[0] != 0.0) p1[i] = p1[i] - c.gamma * c.alpha[0] / c.twiss_beta[0] * (x1[i] - c.twiss_ctr[0]);
    if (c.alpha[1] != 0.0) p2[i] = p2[i] - c.gamma * c.alpha[1] / c.twiss_beta[1] * (x2[i] - c.twiss_ctr[1]);
    double *xs[2] = {x1, x2};
    for (int d = 0; d < 2; d++) {
        if (c.offset_n[d] <= 0) continue;
        const double t = x3[i] - c.offset_xi0[d];
        double y = 0.0, tp = 1.0;
        for (int k = 0; k < c.offset_n[d]; k++) {
            y = y + c.offset_c[d][k] * tp;
            tp = tp * t;
        }
        xs[d][i] = xs[d][i] + y;
    }
}

static int lattice_args(qpad_ctx *c, const qpad_beam_std_cfg &cf, LatticeArgs *a)
{
    memset((void *)a, 0, sizeof(*a));
    a->cf = cf;
    a->dr = c->g.dr;
    a->dz = c->g.dxi;
    const double pi = 4.0 * atan(1.0);
    a->dtheta = 2.0 * pi / (double)cf.num_theta;
    a->coef = copysign(1.0, cf.qm) / ((double)(cf.ppc[0] * cf.ppc[1] * cf.ppc[2]) * (double)cf.num_theta);
    // the radial cells the ranges allow: r <= rmax
    double rmax;
    if (cf.geometry == QPAD_GEOM_CYLINDRICAL) {
        rmax = cf.range_hi[0];
    } else {
        const double mx = fmax(fabs(cf.range_lo[0]), fabs(cf.range_hi[0]));
        const double my = fmax(fabs(cf.range_lo[1]), fabs(cf.range_hi[1]));
        rmax = sqrt(mx * mx + my * my);
    }
    const double imax = floor(rmax / c->g.dr) + 2.0;
    a->ni = (int)fmin((double)c->g.nr, fmax(imax, 0.0));
    const double k0 = floor(cf.range_lo[2] / c->g.dxi) - 1.0, k1 = floor(cf.range_hi[2] / c->g.dxi) + 2.0;
    const int kk0 = (int)fmax(1.0, fmin(k0, (double)c->g.nxi + 1.0));
    const int kk1 = (int)fmax(0.0, fmin(k1, (double)c->g.nxi));
    a->k0 = kk0;
    const int64_t nk = kk1 >= kk0 ? (int64_t)(kk1 - kk0 + 1) : 0;
    a->n = nk * cf.num_theta * a->ni * cf.ppc[0] * cf.ppc[1] * cf.ppc[2];
    if (a->ni <= 0) a->n = 0;
    return 0;
}

// the lattice box of a tabulated density (k0, nk, ni) in place of the one
// the ranges give
static void lattice_box(LatticeArgs *a, const LatticeTable *t, const qpad_beam_std_cfg &cf)
{
    if (!t) return;
    a->den_tab = t->den;
    a->k0 = t->k0;
    a->ni = t->ni;
    a->n = (int64_t)t->nk * cf.num_theta * t->ni * cf.ppc[0] * cf.ppc[1] * cf.ppc[2];
}

int lattice_range_box(qpad_ctx *c, const qpad_beam_std_cfg &cf, int *k0, int *nk, int *ni)
{
    LatticeArgs a;
    QPAD_TRY(lattice_args(c, cf, &a));
    const int64_t per_k = (int64_t)cf.num_theta * a.ni * cf.ppc[0] * cf.ppc[1] * cf.ppc[2];
    *k0 = a.k0;
    *ni = a.ni;
    *nk = per_k > 0 ? (int)(a.n / per_k) : 0;
    return 0;
}

int lattice_count(qpad_ctx *c, const qpad_beam_std_cfg &cf, LatticePlan *pl, const LatticeTable *tab)
{
    memset((void *)pl, 0, sizeof(*pl));
    LatticeArgs a;
    QPAD_TRY(lattice_args(c, cf, &a));
    lattice_box(&a, tab, cf);
    pl->n = a.n;
    if (a.n == 0) return 0;
    if (a.n >= ((int64_t)1 << 31)) {
        set_error("lattice beam: more than 2^31 lattice points inside the ranges");
        return QPAD_E_INVALID;
    }
    QPAD_HIP_CHECK(hipMalloc((void **)&pl->flag, sizeof(uint32_t) * (size_t)a.n));
    QPAD_HIP_CHECK(hipMalloc((void **)&pl->pos, sizeof(uint32_t) * (size_t)a.n));
    size_t bytes = 0;
    QPAD_HIP_CHECK(rocprim::exclusive_scan(nullptr, bytes, pl->flag, pl->pos, 0u, (size_t)a.n,
                                           rocprim::plus<uint32_t>()));
    pl->tmp_bytes = bytes + 256;
    QPAD_HIP_CHECK(hipMalloc(&pl->tmp, pl->tmp_bytes));
    a.flag = pl->flag;
    const unsigned nb = (unsigned)((a.n + 255) / 256);
    hipLaunchKernelGGL(k_lattice_flag, dim3(nb), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    bytes = pl->tmp_bytes;
    QPAD_HIP_CHECK(rocprim::exclusive_scan(pl->tmp, bytes, pl->flag, pl->pos, 0u, (size_t)a.n,
                                           rocprim::plus<uint32_t>(), c->stream));
    uint32_t last_pos = 0, last_flag = 0;
    QPAD_HIP_CHECK(hipMemcpyAsync(&last_pos, pl->pos + (a.n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    QPAD_HIP_CHECK(hipMemcpyAsync(&last_flag, pl->flag + (a.n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    pl->kept = (int64_t)last_pos + last_flag;
    return 0;
}

int lattice_write(qpad_ctx *c, const qpad_beam_std_cfg &cf, const LatticePlan &pl, Beam &bm, const LatticeTable *tab)
{
    int64_t nk = pl.kept;
    if (nk > 0) {
        LatticeArgs a;
        QPAD_TRY(lattice_args(c, cf, &a));
        lattice_box(&a, tab, cf);
        a.flag = pl.flag;
        a.pos = pl.pos;
        a.x1 = bm.x1; a.x2 = bm.x2; a.x3 = bm.x3;
        a.p1 = bm.p1; a.p2 = bm.p2; a.p3 = bm.p3; a.q = bm.q;
        hipLaunchKernelGGL(k_lattice_write, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, c->stream, a);
        QPAD_HIP_CHECK(hipGetLastError());
        if (cf.quiet) {
            hipLaunchKernelGGL(k_beam_quiet, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, c->stream,
                               bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q, nk);
            QPAD_HIP_CHECK(hipGetLastError());
            nk *= 2;
        }
        hipLaunchKernelGGL(k_lattice_shape, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, c->stream, cf, nk,
                           bm.x1, bm.x2, bm.x3, bm.p1, bm.p2);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    bm.npp = nk;
    return 0;
}

void lattice_free(LatticePlan *pl)
{
    if (pl->flag) (void)hipFree(pl->flag);
    if (pl->pos) (void)hipFree(pl->pos);
    if (pl->tmp) (void)hipFree(pl->tmp);
    memset((void *)pl, 0, sizeof(*pl));
}

// ---------------------------------------------------------------------------
// sort by (xi, r) cell: key = mm * nr + nn (0-based cell indices); particles
// outside the box get key = nr * nxi and fall past cell_start[ncell]
// ---------------------------------------------------------------------------
__global__ void k_beam_keys(const double *x1, const double *x2, const double *x3, uint32_t *key, uint32_t *idx,
                            int64_t n, double idr, double idz, double edge_r, double edge_z, int nr, int nxi)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double r = bradius(x1[p], x2[p]);
    double z = x3[p];
    uint32_t k;
    if (!(r < edge_r) || !(z < edge_z) || !(z >= 0.0)) {
        k = (uint32_t)nr * (uint32_t)nxi;
    } else {
        int nn = (int)__dmul_rn(r, idr);
        int mm = (int)(z * idz);
        if (nn > nr - 1) nn = nr - 1;
        if (mm > nxi - 1) mm = nxi - 1;
        k = (uint32_t)mm * (uint32_t)nr + (uint32_t)nn;
    }
    key[p] = k;
    idx[p] = (uint32_t)p;
}

int launch_beam_keys(qpad_ctx *c, int ib)
{
    Beam &bm = c->bm[ib];
    if (bm.npp == 0) return 0;
    hipLaunchKernelGGL(k_beam_keys, dim3((unsigned)((bm.npp + 255) / 256)), dim3(256), 0, c->stream, bm.x1,
                       bm.x2, bm.x3, bm.key, bm.idx, bm.npp, 1.0 / c->g.dr, 1.0 / c->g.dxi,
                       (double)c->g.nr * c->g.dr, (double)c->g.nxi * c->g.dxi, c->g.nr, c->g.nxi);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

struct BPermArgs {
    int64_t n;
    const uint32_t *idx;
    const double *s[7];
    double *d[7];
};

__global__ void k_beam_permute(BPermArgs a)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    uint32_t src = a.idx[p];
    // the seven loads in flight together, then the stores
    double v[7];
#pragma unroll
    for (int f = 0; f < 7; f++) v[f] = a.s[f][src];
#pragma unroll
    for (int f = 0; f < 7; f++) a.d[f][p] = v[f];
}

int launch_beam_sort(qpad_ctx *c, int ib)
{
    Beam &bm = c->bm[ib];
    int64_t n = bm.npp;
    int ncell = c->g.nr * c->g.nxi;
    int bits = 1;
    while ((1u << bits) <= (uint32_t)ncell) bits++;
    if (n > 0) {
        QPAD_TRY(launch_beam_keys(c, ib));
        size_t bytes = bm.sort_tmp_bytes;
        QPAD_HIP_CHECK(rocprim::radix_sort_pairs<SortConfig>(bm.sort_tmp, bytes, bm.key, bm.key_alt, bm.idx, bm.idx_alt,
                                                 (size_t)n, 0, bits, c->stream));
        BPermArgs a;
        a.n = n;
        a.idx = bm.idx_alt;
        const double *src[7] = {bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q};
        double *dst[7] = {bm.a_x1, bm.a_x2, bm.a_x3, bm.a_p1, bm.a_p2, bm.a_p3, bm.a_q};
        for (int f = 0; f < 7; f++) {
            a.s[f] = src[f];
            a.d[f] = dst[f];
        }
        hipLaunchKernelGGL(k_beam_permute, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
        QPAD_HIP_CHECK(hipGetLastError());
        std::swap(bm.x1, bm.a_x1); std::swap(bm.x2, bm.a_x2); std::swap(bm.x3, bm.a_x3);
        std::swap(bm.p1, bm.a_p1); std::swap(bm.p2, bm.a_p2); std::swap(bm.p3, bm.a_p3);
        std::swap(bm.q, bm.a_q);
        std::swap(bm.key, bm.key_alt);
    }
    // cs[ncell] = number of particles inside the box, cs[ncell + 1] = n
    return launch_cell_start(c, bm.key, n, ncell + 2, bm.cell_start);
}

// ---------------------------------------------------------------------------
// charge deposit: qdeposit_part3d (part3d_class.f03:221-356), node-centric
// ---------------------------------------------------------------------------
struct BDepArgs {
    int nr, nxi, nm, ld, M, nbeams;
    double idr, idz;
    const double *x1[kMaxBeams], *x2[kMaxBeams], *x3[kMaxBeams], *q[kMaxBeams];
    const int32_t *cs[kMaxBeams];
    double *h;   // q_beam history [s][nm][ld]
};

__global__ __launch_bounds__(256) void k_beam_qdeposit(BDepArgs a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = a.nr, nxi = a.nxi, nm = a.nm, ld = a.ld;
    const int64_t nnode = (int64_t)(nr + 2) * (nxi + 1);
    if (tid >= nnode) return;
    const int i = (int)(tid % (nr + 2));
    const int s = (int)(tid / (nr + 2)) + 1;   // history slice 1..nxi+1
    double tot[kMaxNm];
    for (int k = 0; k < nm; k++) tot[k] = 0.0;
    for (int b = 0; b < a.nbeams; b++) {
        double acc[kMaxNm];
        for (int k = 0; k < nm; k++) acc[k] = 0.0;
        for (int dk = 0; dk < 2; dk++) {          // xi cell mm = s-1-dk
            const int mm = s - 1 - dk;
            if (mm < 0 || mm >= nxi) continue;
            for (int dj = 0; dj < 2; dj++) {      // r cell nn = i-1-dj
                const int nn = i - 1 - dj;
                if (nn < 0 || nn >= nr) continue;
                const int cell = mm * nr + nn;
                const int q0 = a.cs[b][cell], q1 = a.cs[b][cell + 1];
                for (int p = q0; p < q1; p++) {
                    const double x1 = a.x1[b][p], x2 = a.x2[b][p];
                    double pr = __dmul_rn(bradius(x1, x2), a.idr);
                    double pz = a.x3[b][p] * a.idz;
                    double ph0r = x1 / pr * a.idr, ph0i = -x2 / pr * a.idr;
                    double wr = pr - (double)nn, wz = pz - (double)mm;
                    double wrj = dj ? wr : 1.0 - wr;
                    double wzk = dk ? wz : 1.0 - wz;
                    double w = wrj * wzk;
                    double phr = 1.0 * a.q[b][p], phi = 0.0;
                    acc[0] += w * phr;
                    for (int m = 1; m <= a.M; m++) {
                        double tr = phr * ph0r - phi * ph0i, ti = phr * ph0i + phi * ph0r;
                        phr = tr;
                        phi = ti;
                        acc[2 * m - 1] += w * phr;
                        acc[2 * m] += w * phi;
                    }
                }
            }
        }
        // axis and 1/r for slices 1..nzp (part3d_class.f03:315-351)
        for (int k = 0; k < nm; k++) {
            double v = acc[k];
            if (s <= nxi) {
                if (i == 0) v = 0.0;
                else if (i == 1) v = (k == 0) ? 8.0 * v : 0.0;
                else v = v * (1.0 / (double)(i - 1));
            }
            tot[k] = tot[k] + v;
        }
    }
    for (int k = 0; k < nm; k++) a.h[(size_t)s * nm * ld + (size_t)k * ld + i] = tot[k];
}

int launch_beam_qdeposit(qpad_ctx *c)
{
    BDepArgs a;
    a.nr = c->g.nr;
    a.nxi = c->g.nxi;
    a.nm = c->g.nm;
    a.ld = c->g.ld;
    a.M = c->g.M;
    a.nbeams = c->nbeams;
    a.idr = 1.0 / c->g.dr;
    a.idz = 1.0 / c->g.dxi;
    for (int b = 0; b < kMaxBeams; b++) {
        bool on = b < c->nbeams;
        a.x1[b] = on ? c->bm[b].x1 : nullptr;
        a.x2[b] = on ? c->bm[b].x2 : nullptr;
        a.x3[b] = on ? c->bm[b].x3 : nullptr;
        a.q[b] = on ? c->bm[b].q : nullptr;
        a.cs[b] = on ? c->bm[b].cell_start : nullptr;
    }
    a.h = c->h_q_beam;
    int64_t nnode = (int64_t)(c->g.nr + 2) * (c->g.nxi + 1);
    hipLaunchKernelGGL(k_beam_qdeposit, dim3((unsigned)((nnode + 255) / 256)), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// reduced pusher: push_reduced_part3d (part3d_class.f03:477-576) with the
// bilinear mode gather of interp_emf_part3d (:691-790)
// ---------------------------------------------------------------------------
struct BPushArgs {
    int nr, nxi, nm, ld, M;
    int64_t n;
    double idr, idz, qtmh, dt;
    const double *he, *hb;
    double *x1, *x2, *x3, *p1, *p2, *p3;
    const int32_t *cs;
    int ncell;
};

__global__ __launch_bounds__(256) void k_beam_push(BPushArgs a)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    if (p >= a.cs[a.ncell]) return;   // already outside the box
    const int nm = a.nm, ld = a.ld;
    const size_t sl = (size_t)3 * nm * ld;
    double x1 = a.x1[p], x2 = a.x2[p], x3 = a.x3[p];
    double pr = __dmul_rn(bradius(x1, x2), a.idr);
    double pz = x3 * a.idz;
    double cc = x1 / pr * a.idr, ss = x2 / pr * a.idr;
    int nn = (int)pr, mm = (int)pz;
    pr = pr - (double)nn;
    pz = pz - (double)mm;
    nn = nn + 1;
    mm = mm + 1;
    if (nn > a.nr) nn = a.nr;
    if (mm > a.nxi) mm = a.nxi;
    double wr[2] = {1.0 - pr, pr}, wz[2] = {1.0 - pz, pz};
    double ep[3] = {0.0, 0.0, 0.0}, bp[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++) {
            double wt = wr[j] * wz[k];
            const size_t o = (size_t)(mm + k) * sl + (size_t)(nn + j);
            for (int c3 = 0; c3 < 3; c3++) {
                ep[c3] = ep[c3] + a.he[o + (size_t)c3 * nm * ld] * wt;
                bp[c3] = bp[c3] + a.hb[o + (size_t)c3 * nm * ld] * wt;
            }
        }
    double phr = 1.0, phi = 0.0;
    for (int m = 1; m <= a.M; m++) {
        double tr = phr * cc - phi * ss, ti = phr * ss + phi * cc;
        phr = tr;
        phi = ti;
        const double ph_r = 2.0 * phr, ph_i = 2.0 * phi;
        const int kr = 2 * m - 1, ki = 2 * m;
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 2; j++) {
                double wt = wr[j] * wz[k];
                const size_t o = (size_t)(mm + k) * sl + (size_t)(nn + j);
                for (int c3 = 0; c3 < 3; c3++) {
                    const size_t ocr = o + ((size_t)c3 * nm + kr) * ld, oci = o + ((size_t)c3 * nm + ki) * ld;
                    ep[c3] = ep[c3] + (a.he[ocr] * ph_r - a.he[oci] * ph_i) * wt;
                    bp[c3] = bp[c3] + (a.hb[ocr] * ph_r - a.hb[oci] * ph_i) * wt;
                }
            }
    }
    double t = ep[0] * cc - ep[1] * ss, u = ep[0] * ss + ep[1] * cc;
    ep[0] = t;
    ep[1] = u;
    t = bp[0] * cc - bp[1] * ss;
    u = bp[0] * ss + bp[1] * cc;
    bp[0] = t;
    bp[1] = u;
    for (int c3 = 0; c3 < 3; c3++) {
        ep[c3] = ep[c3] * a.qtmh;
        bp[c3] = bp[c3] * a.qtmh;
    }
    double wp0 = ep[0] - bp[1], wp1 = ep[1] + bp[0], wp2 = ep[2];
    double q1 = a.p1[p], q2 = a.p2[p], q3 = a.p3[p];
    q1 = q1 + wp0; q2 = q2 + wp1; q3 = q3 + wp2;
    q1 = q1 + wp0; q2 = q2 + wp1; q3 = q3 + wp2;
    double dt_gam = a.dt / sqrt(1.0 + q1 * q1 + q2 * q2 + q3 * q3);
    a.p1[p] = q1;
    a.p2[p] = q2;
    a.p3[p] = q3;
    a.x1[p] = x1 + q1 * dt_gam;
    a.x2[p] = x2 + q2 * dt_gam;
    a.x3[p] = x3 - q3 * dt_gam + a.dt;
}

int launch_beam_push(qpad_ctx *c, int ib)
{
    Beam &bm = c->bm[ib];
    if (!bm.cfg.evol || bm.npp == 0) return 0;
    BPushArgs a;
    a.nr = c->g.nr;
    a.nxi = c->g.nxi;
    a.nm = c->g.nm;
    a.ld = c->g.ld;
    a.M = c->g.M;
    a.n = bm.npp;
    a.idr = 1.0 / c->g.dr;
    a.idz = 1.0 / c->g.dxi;
    a.qtmh = bm.qbm * c->g.dt * 0.5;
    a.dt = c->g.dt;
    a.he = c->h_e;
    a.hb = c->h_b;
    a.x1 = bm.x1; a.x2 = bm.x2; a.x3 = bm.x3;
    a.p1 = bm.p1; a.p2 = bm.p2; a.p3 = bm.p3;
    a.cs = bm.cell_start;
    a.ncell = c->g.nr * c->g.nxi;
    hipLaunchKernelGGL(k_beam_push, dim3((unsigned)((bm.npp + 255) / 256)), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace qpad
