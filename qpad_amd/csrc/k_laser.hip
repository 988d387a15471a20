// k_laser.hip -- the laser envelope on the plasma slice (PGC pushers): the
// envelope and its gradients of slice j from the envelope history, and the
// plasma susceptibility chi deposited for the laser solver.
//
// The envelope history is QPAD's field_laser f2 (cfr, cfi: the real and
// imaginary part of the complex envelope), slices -1 .. nxi+1 in HBM as
// [slice + 1][k][ld]; the laser solver that advances it between 3D steps
// (field_laser%solve, sim_lasers%advance) stays with the caller, which
// hands it in through qpad_set_laser and reads chi back per slice.
#include <algorithm>

#include "kernels.h"

namespace qpad {

// ---------------------------------------------------------------------------
// laser_all of slice j (simulation_class.f03:362-367) with one laser:
// zero_field_laser(only_f1) (field_laser_class.f03:952-1012), copy_slice of
// the envelope, set_grad_field_laser (:637-750) and gather (:929-946).
// Thread per (envelope part t, mode array k, node i); every node of the
// outputs is written (nodes that set_grad leaves alone are the zeros of
// zero_field_laser).  Gradient components (r, phi, z): central differences
// in r at nodes 2..nr, the mode's (1/r) d/dphi (-m Im, m Re) / r there, the
// three-point backward difference over slices j, j-1, j-2 in xi at nodes
// 1..nr.  For odd m the axis values 2 a(node 2) / (2 dr) and m times them
// are written at node nr + 1, where set_grad's loop index stands after its
// loop; node 1 keeps zero.  Restated as is.
// ---------------------------------------------------------------------------
struct LaserSliceArgs {
    int nr, nm, ld, j;
    double dr, idrh, idzh;
    const double *h_ar, *h_ai;   // [(nxi + 3)][nm][ld], row s + 1 = slice s
    double *ar, *ai;             // [nm][ld]
    double *arg, *aig;           // [3][nm][ld]
};

__global__ __launch_bounds__(256) void k_laser_slice(const LaserSliceArgs a)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)a.nm * a.ld;
    if (tid >= 2 * per) return;
    const int t = (int)(tid / per);
    const int k = (int)((tid % per) / a.ld), i = (int)(tid % a.ld);
    const int m = k == 0 ? 0 : (k + 1) / 2;
    const double *h = t ? a.h_ai : a.h_ar;
    auto H = [&](int s, int kk, int ii) { return h[((size_t)(s + 1) * a.nm + kk) * a.ld + ii]; };
    const int j = a.j, nr = a.nr;
    (t ? a.ai : a.ar)[(size_t)k * a.ld + i] = H(j, k, i);
    double gr = 0.0, gp = 0.0, gz = 0.0;
    if (i >= 1 && i <= nr) gz = a.idzh * (3.0 * H(j, k, i) - 4.0 * H(j - 1, k, i) + H(j - 2, k, i));
    if (i >= 2 && i <= nr) {
        gr = a.idrh * (H(j, k, i + 1) - H(j, k, i - 1));
        if (m > 0) {
            const double ir = 1.0 / ((double)(i - 1) * a.dr);
            gp = (k == kre(m)) ? -ir * m * H(j, kim(m), i) : ir * m * H(j, kre(m), i);
        }
    }
    if (i == nr + 1 && (m & 1)) {
        const double d_re = 2.0 * a.idrh * H(j, kre(m), 2), d_im = 2.0 * a.idrh * H(j, kim(m), 2);
        gr = k == kre(m) ? d_re : d_im;
        gp = k == kre(m) ? -m * d_im : m * d_re;
    }
    double *g = t ? a.aig : a.arg;
    g[((size_t)0 * a.nm + k) * a.ld + i] = gr;
    g[((size_t)1 * a.nm + k) * a.ld + i] = gp;
    g[((size_t)2 * a.nm + k) * a.ld + i] = gz;
}

int launch_laser_slice(qpad_ctx *c, int j)
{
    LaserSliceArgs a;
    a.nr = c->g.nr;
    a.nm = c->g.nm;
    a.ld = c->g.ld;
    a.j = j;
    a.dr = c->g.dr;
    a.idrh = 0.5 / c->g.dr;
    a.idzh = 0.5 / c->g.dxi;
    a.h_ar = c->lz_ar;
    a.h_ai = c->lz_ai;
    a.ar = c->f[QPAD_F_AR];
    a.ai = c->f[QPAD_F_AI];
    a.arg = c->f[QPAD_F_AR_GRAD];
    a.aig = c->f[QPAD_F_AI_GRAD];
    const int64_t n = 2 * (int64_t)a.nm * a.ld;
    hipLaunchKernelGGL(k_laser_slice, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// deposit_chi_part2d's rescaling (part2d_class.f03:432-455) of the running
// sum: chi = R(chi + deposit of species s), R = node 0 zero, the m = 0 axis
// node times get_deposit_ax_corr(ppc_r) = 12 ppc^2 / (1 + 2 ppc^2)
// (:2581-2586), m > 0 axis zero, node i divided by i - 1.
// sim_lasers%deposit_chi (sim_lasers_class.f03:176-200) calls the deposit
// once per species into one array, so a species' rescaling also acts on the
// species before it; restated as is.  Thread per (mode array k, node i).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_combine_chi(const double *P0, const double *P1, const double *O0,
                                                    const double *O1, const int32_t *os, double *chi, int nr,
                                                    int nm, int ld, double corr)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nm * (nr + 2)) return;
    const int i = (int)(tid % (nr + 2)), k = (int)(tid / (nr + 2));
    double v = 0.0;
    if (i >= 1 && i <= nr) v += cell_partial(P0, O0, os, i - 1, nm, k);
    if (i >= 2) v += cell_partial(P1, O1, os, i - 2, nm, k);
    double *o = chi + (size_t)k * ld + i;
    double x = *o + v;
    if (i == 0) x = 0.0;
    else if (i == 1) x = k == 0 ? x * corr : 0.0;
    else x = x * (1.0 / (double)(i - 1));
    *o = x;
}

// sim_lasers%deposit_chi of slice j: chi = 0, each species' deposit with
// the rescaling, then chi copied into history slice j (copy_slice 1to2)
int launch_deposit_chi(qpad_ctx *c, int j)
{
    const Grid &g = c->g;
    double *chi = c->f[QPAD_F_CHI];
    const size_t n1 = (size_t)g.nm * g.ld;
    QPAD_TRY(launch_fill(c, chi, n1, 0.0));
    for (int s = 0; s < c->nspecies; s++) {
        Species &sp = c->sp[s];
        if (sp.neutral) continue;   // sim_lasers%deposit_chi takes the plasma species only
        const DepRows rows{sp.X0, sp.X1, sp.XO0, sp.XO1};
        QPAD_TRY(launch_chi_deposit_rows(c, s, rows));
        const double ppc = (double)sp.cfg.ppc1;
        const double corr = (12.0 * ppc * ppc) / (1.0 + 2.0 * ppc * ppc);
        const int64_t n = (int64_t)g.nm * (g.nr + 2);
        hipLaunchKernelGGL(k_combine_chi, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, sp.X0, sp.X1,
                           sp.XO0, sp.XO1, sp.ov_start, chi, g.nr, g.nm, g.ld, corr);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    return launch_copy(c, c->h_chi + (size_t)j * n1, chi, n1);
}

}  // namespace qpad
