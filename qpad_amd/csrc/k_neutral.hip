// k_neutral.hip -- a neutral gas ionised by the slice's field
// (neutral_class.f03): the fractions of its ionisation levels per (theta
// bin, radial cell), the electrons each level change releases (a plasma
// species of their own, deposited and pushed like any other), and the ion
// charge they leave behind, which stays in the species' background qn
// (rho_ion) and enters q_spe after the electrons' charge.
//
// Per slice j, in QPAD's order (simulation_class.f03:351-353, 443-447):
//   start: the ion charge of the electrons released at the end of slice
//          j - 1 deposited and added to rho_ion (ion_deposit_neutral);
//   end:   the levels advanced with the ADK rates of the slice's final E,
//          the released electrons appended at rest, then the push.
// Cell c = k * nr + j0 (theta bin k, radial cell j0): QPAD's loop order,
// which also orders the released electrons.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "kernels.h"

namespace qpad {

#define FX(p, c, k, i) (p)[((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)]

// ---------------------------------------------------------------------------
// ionize_neutral (neutral_class.f03:600-753): thread per cell.  |E| at the
// cell centre (the mean of its two nodes, the modes summed at theta_k; for
// m > 0 QPAD reads the real part where it means the imaginary one,
// e_im => e%rf_re(m), restated as is) in GV/m via wp 1.708e-12; above
// 1e-6 and where the cell is not fully ionised, the ADK rate of each level
// and the second-order update of the level fractions with QPAD's overshoot
// rules; the total ion charge rounded to the particles per cell it stands
// for.  The level arithmetic is written without contractions (the rounding
// of the total decides how many electrons a cell releases).  nadd[c] = the
// electrons the cell releases; nadd[nc] = 0 pads the scan.
// ---------------------------------------------------------------------------
struct IonizeArgs {
    int nr, nm, ld, nth, L, ppc_tot;
    double dt, wp, dtheta;
    const double *e;
    const double *rate;   // [L][3]
    double *lv;           // [nc][L + 2]
    int32_t *nadd;        // [nc + 1]
};

template <int M>
__global__ __launch_bounds__(256) void k_ionize(const IonizeArgs a)
{
    const int nc = a.nth * a.nr;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nc) return;
    if (c == nc) {
        a.nadd[nc] = 0;
        return;
    }
    const int nm = a.nm, ld = a.ld;
    const int k = c / a.nr, j = c % a.nr + 1;   // node j, j + 1 bound the cell
    const double theta = a.dtheta * (double)k;
    const double cr = cos(theta), ci = sin(theta);
    double e1 = 0.5 * (FX(a.e, 0, 0, j) + FX(a.e, 0, 0, j + 1));
    double e2 = 0.5 * (FX(a.e, 1, 0, j) + FX(a.e, 1, 0, j + 1));
    double e3 = 0.5 * (FX(a.e, 2, 0, j) + FX(a.e, 2, 0, j + 1));
    double phr = 1.0, phi = 0.0;
#pragma unroll
    for (int m = 1; m <= M; m++) {
        const double tr = phr * cr - phi * ci, ti = phr * ci + phi * cr;
        phr = tr;
        phi = ti;
        const int kr = 2 * m - 1;
        const double s1 = FX(a.e, 0, kr, j) + FX(a.e, 0, kr, j + 1);
        const double s2 = FX(a.e, 1, kr, j) + FX(a.e, 1, kr, j + 1);
        const double s3 = FX(a.e, 2, kr, j) + FX(a.e, 2, kr, j + 1);
        e1 = e1 + s1 * phr - s1 * phi;
        e2 = e2 + s2 * phr - s2 * phi;
        e3 = e3 + s3 * phr - s3 * phi;
    }
    const double eij = sqrt(e1 * e1 + e2 * e2 + e3 * e3) * a.wp * 1.708e-12;
    const int L = a.L, in = L, ii = L + 1;
    double *mi = a.lv + (size_t)c * (L + 2);
    const double old = mi[ii];
    int add = 0;
    if (eij > 1.0e-6 && old < (double)L) {
        const double dt = a.dt;
        auto w_of = [&](int i) {
            const double *r = a.rate + 3 * i;
            return r[0] * pow(eij, -r[2]) * exp(-r[1] / eij) / a.wp;
        };
        // (1 + w dt / 2) w dt of a fraction x, written out in QPAD's order
        auto step = [&](double x, double w) {
            return __dmul_rn(__dmul_rn(__dmul_rn(x, w), dt), __dadd_rn(1.0, __dmul_rn(__dmul_rn(0.5, w), dt)));
        };
        bool shoot = false;
        double cons = step(mi[in], w_of(0));
        if (cons > mi[in]) {
            shoot = true;
            cons = mi[in];
        }
        mi[in] = __dsub_rn(mi[in], cons);
        for (int i = 1; i <= L - 1; i++) {
            const double inj = cons;
            double dens_temp = 0.0;
            if (shoot) {
                dens_temp = __dmul_rn(inj, 0.5);
                shoot = false;
            }
            const double avail = __dadd_rn(mi[i - 1], dens_temp);
            cons = step(avail, w_of(i));
            if (cons > avail) {
                shoot = true;
                cons = avail;
            }
            mi[i - 1] = fmin(__dadd_rn(__dsub_rn(mi[i - 1], cons), inj), 1.0);
        }
        mi[L - 1] = fmin(__dadd_rn(mi[L - 1], cons), 1.0);
        double tot = 0.0;
        for (int i = 1; i <= L; i++) tot = __dadd_rn(tot, __dmul_rn((double)i, mi[i - 1]));
        const double units = (double)(int)__dadd_rn(__ddiv_rn(__dmul_rn(tot, (double)a.ppc_tot), (double)L), 0.5);
        const double now = __dmul_rn(__ddiv_rn((double)L, (double)a.ppc_tot), units);
        mi[ii] = now;
        // add_particles (neutral_class.f03:786-789)
        add = (int)__dadd_rn(__dmul_rn(__ddiv_rn(__dsub_rn(now, old), (double)L), (double)a.ppc_tot), 0.5);
        if (add < 0) add = 0;
    }
    a.nadd[c] = add;
}

// ---------------------------------------------------------------------------
// add_particles (neutral_class.f03:755-837): thread per cell, its released
// electrons at rest at (j0 + (i - 1/2) / n) dr, theta_k, charge
// rn density coef, coef = L sign(q) / (ppc num_theta), written after the
// species' live particles (cell_start[nr]) in QPAD's order; ids count the
// electrons released since the renew; part_add (the ion deposit of the next
// slice start) keeps their positions with the opposite charge
// ---------------------------------------------------------------------------
struct AddArgs {
    int nr, nth;
    int64_t npmax;
    double dr, idr, edge, dtheta, density, coef;
    const int32_t *nadd, *noff;
    const int32_t *cs;       // cell_start[nr]: live particles
    const int32_t *released; // electrons released before this slice
    double *x1, *x2, *p1, *p2, *p3, *q, *gamma, *psi;
    int32_t *id;
    uint32_t *key;
    double *ax1, *ax2, *aq;
};

__device__ inline uint32_t neutral_key(double x1, double x2, double idr, double edge, int nr)
{
    const double r = sqrt(__dadd_rn(__dmul_rn(x1, x1), __dmul_rn(x2, x2)));
    if (!(r < edge)) return (uint32_t)nr;
    const int cc = (int)__dmul_rn(r, idr);
    return (uint32_t)(cc < nr ? cc : nr - 1);
}

__global__ __launch_bounds__(256) void k_neutral_add(const AddArgs a)
{
    const int nc = a.nth * a.nr;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int n = a.nadd[c];
    if (n == 0) return;
    const int k = c / a.nr, j0 = c % a.nr;
    const int live = a.cs[a.nr], base = a.noff[c], rel = *a.released;
    const double theta = (double)k * a.dtheta;
    const double cth = cos(theta), sth = sin(theta);
    for (int i = 1; i <= n; i++) {
        const double rn = (double)j0 + ((double)i - 0.5) / (double)n;
        const double x1 = rn * a.dr * cth, x2 = rn * a.dr * sth;
        const double qp = rn * 1.0 * 1.0 * a.density * a.coef;
        const int64_t p = (int64_t)live + base + i - 1;
        if (p >= a.npmax) return;   // never reached: released electrons fit npmax
        a.x1[p] = x1;
        a.x2[p] = x2;
        a.p1[p] = 0.0;
        a.p2[p] = 0.0;
        a.p3[p] = 0.0;
        a.q[p] = qp;
        a.gamma[p] = 1.0;
        a.psi[p] = 0.0;
        a.id[p] = rel + base + i - 1;
        a.key[p] = neutral_key(x1, x2, a.idr, a.edge, a.nr);
        a.ax1[base + i - 1] = x1;
        a.ax2[base + i - 1] = x2;
        a.aq[base + i - 1] = -qp;
    }
}

// the counts of this slice's release: live += added, released += added;
// the report (live before, added) for the host
__global__ void k_neutral_count(int nr, int nc, const int32_t *noff, int32_t *cs, int32_t *state)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int added = noff[nc];
    const int live = cs[nr];
    state[0] = live;
    state[1] = added;
    state[2] = state[2] + added;   // released since the renew
    cs[nr] = live + added;
}

// ---------------------------------------------------------------------------
// ion_deposit_neutral (neutral_class.f03:904-930): thread per (node i, mode
// array kk): the charge deposit (qdeposit_part2d, part2d_class.f03:231-359:
// linear weights, e^{-im phi}, axis x8 at m = 0, node division) of the
// electrons released at the end of the previous slice with the opposite
// charge, gathered from the cells j0 = i - 1 (inner node) and i - 2 (outer
// node) of every theta bin, added to rho_ion
// ---------------------------------------------------------------------------
struct IonDepArgs {
    int nr, nm, ld, nth, M;
    double idr;
    const int32_t *nadd, *noff;
    const double *ax1, *ax2, *aq;
    double *rho;
};

template <int M>
__global__ __launch_bounds__(256) void k_ion_deposit(const IonDepArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int kk = blockIdx.y;
    if (i > a.nr + 1) return;
    const int m = kk == 0 ? 0 : (kk + 1) / 2;
    const bool im = kk > 0 && (kk & 1) == 0;
    double v = 0.0;
    for (int k = 0; k < a.nth; k++) {
        for (int side = 0; side < 2; side++) {
            const int j0 = side == 0 ? i - 1 : i - 2;
            if (j0 < 0 || j0 >= a.nr) continue;
            const int c = k * a.nr + j0;
            const int n = a.nadd[c], b = a.noff[c];
            for (int t = 0; t < n; t++) {
                const double x1 = a.ax1[b + t], x2 = a.ax2[b + t];
                double pos = sqrt(__dadd_rn(__dmul_rn(x1, x1), __dmul_rn(x2, x2))) * a.idr;
                const double ph0r = x1 / pos * a.idr, ph0i = -x2 / pos * a.idr;
                const int nn = (int)floor(pos);
                pos = pos - (double)nn;
                if (nn != j0) continue;   // the cell the deposit uses (always j0 on this lattice)
                const double w = side == 0 ? 1.0 - pos : pos;
                double phr = 1.0 * a.aq[b + t], phi = 0.0 * a.aq[b + t];
                for (int mm = 1; mm <= m; mm++) {
                    const double tr = phr * ph0r - phi * ph0i, ti = phr * ph0i + phi * ph0r;
                    phr = tr;
                    phi = ti;
                }
                v += w * (im ? phi : phr);
            }
        }
    }
    if (i == 0) v = 0.0;
    else if (i == 1) v = kk == 0 ? 8.0 * v : 0.0;
    else v = v * (1.0 / (double)(i - 1));
    double *o = a.rho + (size_t)kk * a.ld + i;
    *o = *o + v;
}

// the release deposited: nothing left for a later ion deposit (QPAD
// empties part_add, neutral_class.f03:926)
__global__ void k_neutral_clear_adds(int nc, int32_t *nadd)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= nc) nadd[c] = 0;
}

// every cell back to the neutral gas, nothing released (renew_neutral,
// neutral_class.f03:839-878)
__global__ void k_neutral_reset(int nc, int L, double *lv, int32_t *nadd, int32_t *state)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) {
        state[0] = 0;
        state[1] = 0;
        state[2] = 0;
    }
    if (c > nc) return;
    nadd[c] = 0;
    if (c == nc) return;
    for (int l = 0; l < L + 2; l++) lv[(size_t)c * (L + 2) + l] = l == L ? 1.0 : 0.0;
}

static int ntheta_cells(const qpad_ctx *c, const Species &sp) { return sp.cfg.num_theta * c->g.nr; }

int neutral_scan_bytes(int nc, size_t *bytes)
{
    int32_t *z = nullptr;
    QPAD_HIP_CHECK(rocprim::exclusive_scan(nullptr, *bytes, z, z, 0, (size_t)nc + 1, rocprim::plus<int32_t>()));
    return 0;
}

int launch_neutral_reset(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    const int nc = ntheta_cells(c, sp);
    hipLaunchKernelGGL(k_neutral_reset, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, c->stream, nc,
                       sp.multi_max, sp.levels, sp.nadd, sp.nstate);
    QPAD_HIP_CHECK(hipGetLastError());
    return launch_fill(c, sp.qn, (size_t)c->g.nm * c->g.ld, 0.0);
}

int launch_ion_deposit(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    IonDepArgs a;
    a.nr = c->g.nr;
    a.nm = c->g.nm;
    a.ld = c->g.ld;
    a.nth = sp.cfg.num_theta;
    a.M = c->g.M;
    a.idr = 1.0 / c->g.dr;
    a.nadd = sp.nadd;
    a.noff = sp.noff;
    a.ax1 = sp.add_x1;
    a.ax2 = sp.add_x2;
    a.aq = sp.add_q;
    a.rho = sp.qn;
    const dim3 grid((unsigned)((c->g.nr + 2 + 255) / 256), (unsigned)c->g.nm), blk(256);
    switch (c->g.M) {
    case 0: hipLaunchKernelGGL(k_ion_deposit<0>, grid, blk, 0, c->stream, a); break;
    case 1: hipLaunchKernelGGL(k_ion_deposit<1>, grid, blk, 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL(k_ion_deposit<2>, grid, blk, 0, c->stream, a); break;
    case 3: hipLaunchKernelGGL(k_ion_deposit<3>, grid, blk, 0, c->stream, a); break;
    default: hipLaunchKernelGGL(k_ion_deposit<4>, grid, blk, 0, c->stream, a); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());
    const int nc = ntheta_cells(c, sp);
    hipLaunchKernelGGL(k_neutral_clear_adds, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, c->stream, nc,
                       sp.nadd);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// the levels and the release of species s from the slice's final E; the
// release's size comes back to the host (it sizes the push and the sort)
int launch_neutral_update(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    const int nc = ntheta_cells(c, sp);
    IonizeArgs a;
    a.nr = c->g.nr;
    a.nm = c->g.nm;
    a.ld = c->g.ld;
    a.nth = sp.cfg.num_theta;
    a.L = sp.multi_max;
    a.ppc_tot = sp.cfg.ppc1 * sp.cfg.ppc2;
    a.dt = c->g.dxi;
    a.wp = sp.wp;
    a.dtheta = 2.0 * (4.0 * atan(1.0)) / (double)sp.cfg.num_theta;
    a.e = c->f[QPAD_F_E];
    a.rate = sp.rate;
    a.lv = sp.levels;
    a.nadd = sp.nadd;
    const dim3 grid((unsigned)((nc + 1 + 255) / 256)), blk(256);
    switch (c->g.M) {
    case 0: hipLaunchKernelGGL(k_ionize<0>, grid, blk, 0, c->stream, a); break;
    case 1: hipLaunchKernelGGL(k_ionize<1>, grid, blk, 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL(k_ionize<2>, grid, blk, 0, c->stream, a); break;
    case 3: hipLaunchKernelGGL(k_ionize<3>, grid, blk, 0, c->stream, a); break;
    default: hipLaunchKernelGGL(k_ionize<4>, grid, blk, 0, c->stream, a); break;
    }
    QPAD_HIP_CHECK(hipGetLastError());
    size_t bytes = sp.scan_bytes;
    QPAD_HIP_CHECK(rocprim::exclusive_scan(sp.scan_tmp, bytes, sp.nadd, sp.noff, 0, (size_t)nc + 1,
                                           rocprim::plus<int32_t>(), c->stream));
    AddArgs d;
    d.nr = c->g.nr;
    d.nth = sp.cfg.num_theta;
    d.npmax = sp.npmax;
    d.dr = c->g.dr;
    d.idr = 1.0 / c->g.dr;
    d.edge = (double)c->g.nr * c->g.dr;
    d.dtheta = a.dtheta;
    d.density = sp.cfg.density;
    d.coef = (double)sp.multi_max * copysign(1.0, sp.cfg.qm) /
             ((double)(sp.cfg.ppc1 * sp.cfg.ppc2) * (double)sp.cfg.num_theta);
    d.nadd = sp.nadd;
    d.noff = sp.noff;
    d.cs = sp.cell_start;
    d.released = sp.nstate + 2;
    d.x1 = sp.x1; d.x2 = sp.x2; d.p1 = sp.p1; d.p2 = sp.p2; d.p3 = sp.p3;
    d.q = sp.q; d.gamma = sp.gamma; d.psi = sp.psi; d.id = sp.id; d.key = sp.key;
    d.ax1 = sp.add_x1;
    d.ax2 = sp.add_x2;
    d.aq = sp.add_q;
    hipLaunchKernelGGL(k_neutral_add, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, d);
    QPAD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_neutral_count, dim3(1), dim3(64), 0, c->stream, c->g.nr, nc, sp.noff, sp.cell_start,
                       sp.nstate);
    QPAD_HIP_CHECK(hipGetLastError());
    QPAD_HIP_CHECK(hipMemcpyAsync(sp.nstate_host, sp.nstate, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    sp.npp = std::max<int64_t>(sp.npp, (int64_t)sp.nstate_host[0] + sp.nstate_host[1]);
    return 0;
}

// ---------------------------------------------------------------------------
// ADK coefficients per level, w = C1 E^-C3 exp(-C2 / E) (E in GV/m, w in
// 1/s; neutral_class.f03:23-299, the elements of atomic number 1, 2, 3, 6,
// 7, 8, 10, 18).  Row i is the tunnelling rate (l = 0, m = 0) of charge
// state i with its ionisation energy U: n* = i sqrt(U_H / U), C3 = 2n* - 1,
// C2 = (2/3) (U/U_H)^(3/2) E_a, C1 = w_a 2^(2n*) / (n* Gamma(2n*))
// (U/2U_H) (2 (U/U_H)^(3/2) E_a)^(2n* - 1).
// ---------------------------------------------------------------------------
static const double kAdkH[1][3] = {{8.522542995398661e19, 342.53947239007687, 1.0005337056631487}};
static const double kAdkHe[2][3] = {{7.2207661763501e18, 832.809878216992, 0.48776427204592254},
                                    {2.7226733893691e21, 2742.1316798375965, 1.0000920088118899}};
static const double kAdkLi[3][3] = {{3.460272990838495e21, 85.51998980232813, 2.1770706138013733},
                                    {3.6365138642921554e20, 4493.713340713575, 0.6964625952167312},
                                    {2.0659396971422902e22, 9256.32561931876, 0.9999745128918196}};
static const double kAdkC[6][3] = {{1.8809310466441196e20, 258.1083104918942, 1.1984440550739555},
                                   {5.509352426405289e23, 822.523017116882, 1.987881643259267},
                                   {4.572825176241475e25, 2263.6763526780164, 2.198152910400111},
                                   {9.835564484116034e27, 3537.9525592618193, 2.674447703635764},
                                   {7.536123777733801e22, 53034.28979013096, 0.8628040191262976},
                                   {6.587972799911127e23, 74090.46603762524, 0.9996157827449552}};
static const double kAdkN[7][3] = {{6.441568477283914e19, 378.4956025896395, 0.9350660865589433},
                                   {1.4701035287257558e23, 1100.1258100528623, 1.711847679118431},
                                   {4.963404971976876e25, 2232.374603536868, 2.2130314611564854},
                                   {1.429510866064162e27, 4658.081063546513, 2.3525381223431663},
                                   {1.3012849307544791e29, 6615.849778853221, 2.7281285044063783},
                                   {2.1409091975134166e23, 88606.4475724123, 0.8838466900060211},
                                   {1.4212916247536375e24, 117682.27130021283, 0.9994495038042972}};
static const double kAdkO[8][3] = {{8.471015773053202e19, 343.2810702373722, 0.9990920676510977},
                                   {4.659951179950614e22, 1421.770921017062, 1.4896379815554006},
                                   {1.375943467286961e25, 2781.3610873184325, 1.985966133301126},
                                   {1.4410401568574448e27, 4652.670876621719, 2.353837077475246},
                                   {2.1812887948673846e28, 8303.41171218499, 2.4562134926799617},
                                   {1.0859737168339012e30, 11088.09514588445, 2.766300924086769},
                                   {5.135265231597041e23, 137319.4144041063, 0.8991975448040144},
                                   {2.7649244731504574e24, 175715.8649323391, 0.999259077262769}};
static const double kAdkNe[10][3] = {{1.2380498982127483e19, 684.0495717312216, 0.5886207199036224},
                                     {1.6861743227006676e22, 1790.87087951275, 1.3052851455522987},
                                     {4.012462791033353e24, 3450.2505445538854, 1.7789909006599105},
                                     {1.4251452589514637e26, 6544.999998027671, 1.9932260907792414},
                                     {6.670617838240815e27, 9689.667572133209, 2.282840577038655},
                                     {1.9391656057538128e29, 13557.847041347119, 2.522116807457813},
                                     {1.3881927947789378e30, 20383.797283570835, 2.586898525643846},
                                     {3.1075134088626144e31, 25254.466329809107, 2.8167466774677434},
                                     {2.1020658703689154e24, 282468.0585535283, 0.9200040384519763},
                                     {8.391062111505151e24, 343429.9456106982, 0.9988031600227598}};
static const double kAdkAr[18][3] = {{4.583155091154178e19, 427.3616288942586, 0.858307468844768},
                                     {2.347069522730368e23, 992.0665907024354, 1.8069357289301329},
                                     {1.931637188507353e26, 1775.9423286421918, 2.4675897958480384},
                                     {2.3004409750295724e28, 3141.43462246603, 2.8229627269683206},
                                     {3.468592152886133e30, 4422.602894394151, 3.263766733070865},
                                     {2.965675081118506e32, 5958.158993507227, 3.6326551045811497},
                                     {2.1249901067576396e33, 9479.178143644607, 3.629746649292696},
                                     {8.991061313075681e34, 11737.03857874398, 3.92742350307885},
                                     {7.611312705906995e29, 59342.85460567607, 2.229751529947494},
                                     {6.602179339057654e30, 71781.6325368281, 2.3680482459080237},
                                     {4.815600903407338e31, 85819.59318777094, 2.490706079578982},
                                     {2.055113377686043e32, 105186.65111959413, 2.558310025830369},
                                     {1.1945429165015238e33, 122591.38628416909, 2.663021333332675},
                                     {6.212435859768618e33, 141745.69057711778, 2.7584389081812857},
                                     {1.6060852787486324e34, 170916.68217921766, 2.783373249974532},
                                     {9.327897444360857e34, 190110.7378894357, 2.894937691775023},
                                     {6.6878740700780965e25, 1.8068758775178685e6, 0.9536895901987263},
                                     {1.5247916822996254e26, 2.0115330349032478e6, 0.9959340925415332}};

const double *adk_row(int element, int level)
{
    const double(*t)[3] = nullptr;
    int n = 0;
    switch (element) {
    case 1: t = kAdkH; n = 1; break;
    case 2: t = kAdkHe; n = 2; break;
    case 3: t = kAdkLi; n = 3; break;
    case 6: t = kAdkC; n = 6; break;
    case 7: t = kAdkN; n = 7; break;
    case 8: t = kAdkO; n = 8; break;
    case 10: t = kAdkNe; n = 10; break;
    case 18: t = kAdkAr; n = 18; break;
    default: return nullptr;
    }
    return (level >= 1 && level <= n) ? t[level - 1] : nullptr;
}

}  // namespace qpad
