// qpad_ctx.cpp -- host side of libqpad_hip.so: context lifetime, solver
// matrices, the per-slice / per-3D-step orchestration and the C ABI
// declared in include/qpad_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"

namespace qpad {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define TRY(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != 0) return rc_; \
    } while (0)

// ---------------------------------------------------------------------------
// Tridiagonal matrices of field_solver (source/fields/field_solver_class.f03:
// 256-561): rows 0..nr-1 = grid points 1..nr of one radial partition
// (noff = 0).  k = m for psi/Ez/Bz/beam bt, m+1 for B+, m-1 for B-; B+/B-
// carry -relax_fac on the diagonal; open outer boundary folds
// x_{N+1} = coef * x_N into the diagonal.
// ---------------------------------------------------------------------------
static void build_matrix(int kind, int m, int nr, double dr, double relax, int bnd, double *a, double *b,
                         double *c)
{
    const double dr2 = dr * dr;
    const double m2 = (double)(m * m);
    for (int r = 1; r < nr; r++) {
        const double j = (double)r;
        a[r] = 1.0 - 0.5 / j;
        c[r] = 1.0 + 0.5 / j;
        if (kind == FK_BPLUS) {
            const double t = (double)(m + 1) / j;
            b[r] = -2.0 - t * t - relax;
        } else if (kind == FK_BMINUS) {
            const double t = (double)(m - 1) / j;
            b[r] = -2.0 - t * t - relax;
        } else {
            b[r] = -2.0 - m2 / (j * j);
        }
    }
    bool identity;
    if (kind == FK_BPLUS) {
        identity = true;
    } else if (kind == FK_BMINUS) {
        identity = (m != 1);
        if (!identity) { a[0] = 0.0; b[0] = -4.0 - relax; c[0] = 4.0; }
    } else {
        identity = (m != 0);
        if (!identity) { a[0] = 0.0; b[0] = -4.0; c[0] = 4.0; }
    }
    if (identity) {
        // arbitrary non-singular first row; a(2) = 0 decouples the axis
        a[0] = 0.0; b[0] = 1.0; c[0] = 0.0;
        if (nr > 1) a[1] = 0.0;
    }
    const int rl = nr - 1;
    if (bnd == QPAD_BND_ZERO) {
        c[rl] = 0.0;
    } else {
        const double jmax = (double)nr;
        double coef = 0.0;
        bool fold = true;
        switch (kind) {
        case FK_PSI:
        case FK_EZ:
        case FK_BZ:
            if (m == 0) fold = false;
            else coef = 1.0 - (double)m / jmax;
            break;
        case FK_BT:
            coef = (m == 0) ? 1.0 + 1.0 / (jmax * std::log(jmax * dr)) : 1.0 - (double)m / jmax;
            break;
        default:  // B+ and B- both fold with m+1
            coef = 1.0 - (double)(m + 1) / jmax;
            break;
        }
        if (fold) b[rl] = b[rl] + coef * c[rl];
        c[rl] = 0.0;
    }
    for (int r = 0; r < nr; r++) {
        a[r] = a[r] / dr2;
        b[r] = b[r] / dr2;
        c[r] = c[r] / dr2;
    }
}

static void thomas_factor(int n, const double *a, const double *b, const double *c, double *lo, double *inv,
                          double *cp)
{
    inv[0] = 1.0 / b[0];
    cp[0] = c[0] * inv[0];
    lo[0] = 0.0;
    for (int r = 1; r < n; r++) {
        const double den = b[r] - a[r] * cp[r - 1];
        inv[r] = 1.0 / den;
        cp[r] = c[r] * inv[r];
        lo[r] = a[r];
    }
}

// ---------------------------------------------------------------------------
// event timing of the dominant kernels
// ---------------------------------------------------------------------------
static void pool_drain(qpad_ctx *c, EventPool &p)
{
    if (p.used == 0) return;
    (void)hipEventSynchronize(p.ev[2 * p.used - 1]);
    for (size_t i = 0; i < p.used; i++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]) == hipSuccess) p.acc_ms += ms;
    }
    p.used = 0;
}

void time_begin(qpad_ctx *c, int which, hipEvent_t *stop)
{
    *stop = nullptr;
    if (!c->timing) return;
    EventPool &p = c->pools[which];
    if (p.used * 2 + 2 > p.ev.size()) {
        // grow the pool up to 8192 events; when full, or when an event cannot
        // be created, read out the recorded pairs and reuse them
        const size_t want = std::min<size_t>(8192, p.ev.size() + 512);
        while (p.ev.size() < want) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) break;
            p.ev.push_back(e);
        }
        if (p.used * 2 + 2 > p.ev.size()) pool_drain(c, p);
        if (p.ev.size() < 2) return;
    }
    (void)hipEventRecord(p.ev[2 * p.used], c->stream);
    *stop = p.ev[2 * p.used + 1];
}

void time_end(qpad_ctx *c, int which, hipEvent_t stop)
{
    if (!stop) return;
    EventPool &p = c->pools[which];
    (void)hipEventRecord(stop, c->stream);
    p.used++;
    p.launches++;
}

// ---------------------------------------------------------------------------
// allocation helpers
// ---------------------------------------------------------------------------
template <class T>
static int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(QPAD_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return 0;
}

template <class T>
static void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static int zero_dev(qpad_ctx *c, double *p, size_t n) { return launch_fill(c, p, n, 0.0); }

static size_t f1_size(const qpad_ctx *c, int ncomp) { return (size_t)ncomp * c->g.nm * c->g.ld; }

static int sync(qpad_ctx *c)
{
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

static int refresh_npp(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    int32_t v = 0;
    QPAD_HIP_CHECK(hipMemcpyAsync(&v, sp.cell_start + c->g.nr, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    sp.npp = v;
    return 0;
}

static int refresh_beam_npp(qpad_ctx *c, int b)
{
    Beam &bm = c->bm[b];
    int32_t v = 0;
    size_t ncell = (size_t)c->g.nr * c->g.nxi;
    QPAD_HIP_CHECK(hipMemcpyAsync(&v, bm.cell_start + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    bm.npp = v;
    return 0;
}

// species charge setup after (re)injection: sort, background qn = -q
static int species_charge_setup(qpad_ctx *c, int s)
{
    TRY(launch_sort(c, s));
    TRY(refresh_npp(c, s));
    TRY(launch_qdeposit(c, s));
    if (c->sp[s].cfg.neutralized) TRY(launch_combine_q(c, s, c->sp[s].qn, -1.0, false));
    return 0;
}

static void species_free(Species &sp)
{
    dfree(sp.x1); dfree(sp.x2); dfree(sp.p1); dfree(sp.p2); dfree(sp.p3); dfree(sp.q);
    dfree(sp.gamma); dfree(sp.psi); dfree(sp.id);
    dfree(sp.a_x1); dfree(sp.a_x2); dfree(sp.a_p1); dfree(sp.a_p2); dfree(sp.a_p3); dfree(sp.a_q);
    dfree(sp.a_gamma); dfree(sp.a_psi); dfree(sp.a_id);
    dfree(sp.key); dfree(sp.key_alt);
    dfree(sp.cell_start); dfree(sp.qn); dfree(sp.P0); dfree(sp.P1);
    dfree(sp.ov_start); dfree(sp.ov_desc); dfree(sp.O0); dfree(sp.O1);
    dfree(sp.hist); dfree(sp.tot);
    dfree(sp.qsl); dfree(sp.h_q);
    dfree(sp.X0); dfree(sp.X1); dfree(sp.XO0); dfree(sp.XO1);
    dfree(sp.rate); dfree(sp.levels); dfree(sp.nadd); dfree(sp.noff);
    dfree(sp.add_x1); dfree(sp.add_x2); dfree(sp.add_q); dfree(sp.nstate); dfree(sp.h_rho);
    if (sp.scan_tmp) (void)hipFree(sp.scan_tmp);
    sp.scan_tmp = nullptr;
    dfree(sp.den_perp); dfree(sp.inj_keep); dfree(sp.inj_off);
    if (sp.inj_tmp) (void)hipFree(sp.inj_tmp);
    sp.inj_tmp = nullptr;
    if (sp.inj_host) (void)hipHostFree(sp.inj_host);
    sp.inj_host = nullptr;
    if (sp.nstate_host) (void)hipHostFree(sp.nstate_host);
    sp.nstate_host = nullptr;
}

// a species back to its state at the 3D-step start: the plasma injected
// (and its charge set up), or a neutral gas with nothing ionised
static int species_restart(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    if (!sp.neutral) {
        TRY(launch_inject(c, s));
        return species_charge_setup(c, s);
    }
    sp.npp = 0;
    TRY(launch_neutral_reset(c, s));
    return launch_sort(c, s);   // no particles: every cell start 0
}

// the per-species charge arrays of species s (zeroed)
static int species_q_alloc(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    if (sp.h_q) return 0;
    const size_t n1 = f1_size(c, 1), nh = ((size_t)c->g.nxi + 2) * n1;
    TRY(dalloc(&sp.qsl, n1));
    TRY(dalloc(&sp.h_q, nh));
    TRY(zero_dev(c, sp.qsl, n1));
    return zero_dev(c, sp.h_q, nh);
}

static bool is_pgc(int push_type) { return push_type == QPAD_PUSH2_STD_PGC || push_type == QPAD_PUSH2_ROBUST_PGC; }

// the susceptibility deposit rows of species s (every row a consumer reads
// is written by the deposit, so they need no clearing)
static int species_chi_alloc(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    if (sp.X0) return 0;
    const size_t nm = (size_t)c->g.nm;
    TRY(dalloc(&sp.X0, (size_t)c->g.nr * nm));
    TRY(dalloc(&sp.X1, (size_t)c->g.nr * nm));
    TRY(dalloc(&sp.XO0, (size_t)sp.n_ov_max * nm));
    return dalloc(&sp.XO1, (size_t)sp.n_ov_max * nm);
}

// the neutral state of species s (its buffers freed with the species)
static int neutral_alloc(qpad_ctx *c, int s, const qpad_neutral_cfg &nz)
{
    Species &sp = c->sp[s];
    sp.neutral = true;
    sp.multi_max = nz.ion_max;
    sp.wp = sqrt(nz.n0) * 5.641460231180626e4;   // sim_plasma_class.f03:84
    const int nc = sp.cfg.num_theta * c->g.nr, L = sp.multi_max;
    TRY(dalloc(&sp.rate, (size_t)3 * L));
    TRY(dalloc(&sp.levels, (size_t)nc * (L + 2)));
    TRY(dalloc(&sp.nadd, (size_t)nc + 1));
    TRY(dalloc(&sp.noff, (size_t)nc + 1));
    TRY(dalloc(&sp.add_x1, (size_t)sp.npmax));
    TRY(dalloc(&sp.add_x2, (size_t)sp.npmax));
    TRY(dalloc(&sp.add_q, (size_t)sp.npmax));
    TRY(dalloc(&sp.nstate, 3));
    TRY(dalloc(&sp.h_rho, ((size_t)c->g.nxi + 2) * f1_size(c, 1)));
    TRY(zero_dev(c, sp.h_rho, ((size_t)c->g.nxi + 2) * f1_size(c, 1)));
    TRY(neutral_scan_bytes(nc, &sp.scan_bytes));
    if (hipMalloc(&sp.scan_tmp, std::max<size_t>(sp.scan_bytes, 1)) != hipSuccess)
        return fail(QPAD_E_NOMEM, "hipMalloc of the neutral scan scratch failed");
    if (hipHostMalloc((void **)&sp.nstate_host, 3 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return fail(QPAD_E_NOMEM, "hipHostMalloc failed");
    std::vector<double> r(3 * L);
    for (int i = 0; i < L; i++)
        for (int t = 0; t < 3; t++) r[3 * i + t] = adk_row(nz.element, i + 1)[t];
    QPAD_HIP_CHECK(hipMemcpyAsync(sp.rate, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

// the laser state (a zero envelope history, the chi history, every
// species' chi rows); on failure the context is left without it
static int laser_ensure(qpad_ctx *c)
{
    if (c->laser) return 0;
    const size_t n1 = f1_size(c, 1);
    const size_t nh = ((size_t)c->g.nxi + 3) * n1, nc = ((size_t)c->g.nxi + 2) * n1;
    int rc = dalloc(&c->lz_ar, nh);
    if (!rc) rc = dalloc(&c->lz_ai, nh);
    if (!rc) rc = dalloc(&c->h_chi, nc);
    for (int s = 0; s < c->nspecies && !rc; s++) rc = species_chi_alloc(c, s);
    if (!rc) rc = zero_dev(c, c->lz_ar, nh);
    if (!rc) rc = zero_dev(c, c->lz_ai, nh);
    if (!rc) rc = zero_dev(c, c->h_chi, nc);
    if (!rc) rc = sync(c);
    if (rc) {
        const std::string msg = g_err;
        (void)hipStreamSynchronize(c->stream);
        dfree(c->lz_ar); dfree(c->lz_ai); dfree(c->h_chi);
        g_err = msg;
        return rc;
    }
    c->laser = true;
    return 0;
}

static void beam_free(Beam &bm)
{
    dfree(bm.x1); dfree(bm.x2); dfree(bm.x3); dfree(bm.p1); dfree(bm.p2); dfree(bm.p3); dfree(bm.q);
    dfree(bm.a_x1); dfree(bm.a_x2); dfree(bm.a_x3); dfree(bm.a_p1); dfree(bm.a_p2); dfree(bm.a_p3);
    dfree(bm.a_q);
    dfree(bm.key); dfree(bm.key_alt); dfree(bm.idx); dfree(bm.idx_alt); dfree(bm.cell_start);
    if (bm.sort_tmp) (void)hipFree(bm.sort_tmp);
    bm.sort_tmp = nullptr;
}

}  // namespace qpad

using namespace qpad;

extern "C" {

const char *qpad_last_error(void) { return g_err.c_str(); }

int qpad_abi_version(void) { return QPAD_HIP_ABI_VERSION; }

int qpad_create(const qpad_grid *gr, qpad_ctx **out)
{
    if (!gr || !out) return fail(QPAD_E_INVALID, "null argument");
    *out = nullptr;
    if (gr->nr < 8) return fail(QPAD_E_INVALID, "nr must be >= 8");
    if (gr->nxi < 1) return fail(QPAD_E_INVALID, "nxi must be >= 1");
    if (gr->max_mode < 0 || gr->max_mode > kMaxMode)
        return fail(QPAD_E_INVALID, "max_mode must be in 0..4");
    if (!(gr->dr > 0.0) || !(gr->dxi > 0.0)) return fail(QPAD_E_INVALID, "dr and dxi must be > 0");
    if (gr->bnd != QPAD_BND_OPEN && gr->bnd != QPAD_BND_ZERO)
        return fail(QPAD_E_INVALID, "field_boundary must be open or zero");
    if (gr->iter_max < 1) return fail(QPAD_E_INVALID, "iter_max must be >= 1");
    int dev = 0;
    QPAD_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    QPAD_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    if (gr->nr > 16 * kSolveThreads) return fail(QPAD_E_INVALID, "nr must be <= 8192");
    // the field solves stage (K + 1) * kSolveThreads doubles in LDS
    const int kpt = rows_per_thread(gr->nr);
    size_t lds_need = sizeof(double) * (size_t)(kpt + 1) * kSolveThreads;
    if (lds_need + 1024 > prop.sharedMemPerBlock)
        return fail(QPAD_E_INVALID, "nr too large for this device's LDS");

    if (field_kernels_init(lds_need) != 0 || slice_kernels_init() != 0 || sort_kernels_init() != 0)
        return QPAD_E_HIP;
    TRY(noalloc_init_device(dev));
    qpad_ctx *c = new qpad_ctx{};
    c->device = dev;
    Grid &g = c->g;
    g.nr = gr->nr;
    g.nxi = gr->nxi;
    g.M = gr->max_mode;
    g.nm = 2 * gr->max_mode + 1;
    g.ld = ((gr->nr + 2) + 7) & ~7;
    g.dr = gr->dr;
    g.dxi = gr->dxi;
    g.idr = 1.0 / gr->dr;
    g.relax_fac = gr->relax_fac;
    g.dt = gr->dt;
    g.bnd = gr->bnd;
    g.iter_max = gr->iter_max;
    g.reltol = gr->iter_reltol;
    g.abstol = gr->iter_abstol;
    int rc = 0;
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete c;
        return fail(QPAD_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    static const int ncomp[QPAD_NFIELD] = {1, 3, 3, 3, 3, 3, 3, 2, 3, 2, 1, 1, 1, 1, 3, 3, 1};
    for (int f = 0; f < QPAD_NFIELD && !rc; f++) {
        c->ncomp[f] = ncomp[f];
        rc = dalloc(&c->f[f], f1_size(c, ncomp[f]));
    }
    const size_t ns = (size_t)g.nxi + 2;
    if (!rc) rc = dalloc(&c->h_psi, ns * f1_size(c, 1));
    if (!rc) rc = dalloc(&c->h_e, ns * f1_size(c, 3));
    if (!rc) rc = dalloc(&c->h_b, ns * f1_size(c, 3));
    if (!rc) rc = dalloc(&c->h_b_spe, ns * f1_size(c, 3));
    if (!rc) rc = dalloc(&c->h_e_spe, ns * f1_size(c, 3));
    if (!rc) rc = dalloc(&c->h_cu, ns * f1_size(c, 3));
    if (!rc) rc = dalloc(&c->h_q_spe, ns * f1_size(c, 1));
    if (!rc) rc = dalloc(&c->h_q_beam, ns * f1_size(c, 1));
    if (!rc) rc = dalloc(&c->h_b_beam, ns * f1_size(c, 3));
    if (!rc) rc = dalloc(&c->conv_re, (size_t)g.nr + 2);
    if (!rc) rc = dalloc(&c->conv_im, (size_t)g.nr + 2);
    if (!rc) rc = dalloc(&c->conv_blk, 2 * 2 * (size_t)conv_blocks(g.nr));
    if (!rc) rc = dalloc(&c->it, 1);
    if (!rc) rc = dalloc(&c->sc_ef, 1);
    // factors stored transposed for the solve kernels: [kind][m][lo|inv|cp]
    // then row r = t * K + kk at kk * kSolveThreads + t; rows >= nr are
    // identity rows (lo = 0, inv = 1, cp = 0)
    const size_t kt = (size_t)kpt * kSolveThreads;
    size_t nfac = (size_t)FK_N * (kMaxMode + 1) * 3 * kt;
    if (!rc) rc = dalloc(&c->fac, nfac);
    if (!rc) rc = dalloc(&c->src, (size_t)src_rows_host(g.M, g.nm) * g.nr);
    if (!rc) {
        he = hipHostMalloc((void **)&c->it_host, sizeof(IterState), hipHostMallocDefault);
        if (he != hipSuccess) rc = fail(QPAD_E_NOMEM, "hipHostMalloc failed");
    }
    if (!rc) {
        he = hipHostMalloc((void **)&c->sc_ef_host, sizeof(double), hipHostMallocDefault);
        if (he != hipSuccess) rc = fail(QPAD_E_NOMEM, "hipHostMalloc failed");
    }
    if (!rc) {
        he = hipHostMalloc((void **)&c->it_report, sizeof(IterState), hipHostMallocCoherent);
        if (he != hipSuccess) rc = fail(QPAD_E_NOMEM, "hipHostMalloc (coherent) failed");
        else memset(c->it_report, 0, sizeof(IterState));
    }
    if (!rc) {
        he = hipEventCreateWithFlags(&c->dec_ev, hipEventDisableTiming);
        if (he != hipSuccess) rc = fail(QPAD_E_HIP, "hipEventCreate failed");
    }
    c->adaptive = 1;
    c->spec_margin = 0;
    c->spec_guess = 2;
    c->dt2d = g.dxi;
    c->clamp2d = 0.0;
    c->last_nsub = 1;
    c->last_dsub = g.dxi;
    c->dep_cells = kDepWaves;
    // the matrix form where it measured faster on MI355X: modes 0..3 and up
    // (C3, modes 0..4: 18.6 against 26.0 us per current deposit); the
    // reduce-scatter up to modes 0..2 (C2: 21.5 against 27.5 us)
    c->dep_mfma = gr->max_mode >= 3 ? 1 : 0;
    if (rc) {
        qpad_destroy(c);
        return rc;
    }
    // solver factors
    std::vector<double> hf(nfac, 0.0), a(g.nr), b(g.nr), cc(g.nr), lo(g.nr), inv(g.nr), cp(g.nr);
    for (int kind = 0; kind < FK_N; kind++)
        for (int m = 0; m <= g.M; m++) {
            build_matrix(kind, m, g.nr, g.dr, g.relax_fac, g.bnd, a.data(), b.data(), cc.data());
            thomas_factor(g.nr, a.data(), b.data(), cc.data(), lo.data(), inv.data(), cp.data());
            double *base = hf.data() + (((size_t)kind * (kMaxMode + 1) + m) * 3) * kt;
            for (size_t r = 0; r < kt; r++) {
                const size_t t = r / kpt, kk = r % kpt, at = kk * kSolveThreads + t;
                const bool in = r < (size_t)g.nr;
                base[at] = in ? lo[r] : 0.0;
                base[kt + at] = in ? inv[r] : 1.0;
                base[2 * kt + at] = in ? cp[r] : 0.0;
            }
        }
    rc = 0;
    he = hipMemcpy(c->fac, hf.data(), nfac * sizeof(double), hipMemcpyHostToDevice);
    if (he != hipSuccess) rc = fail(QPAD_E_HIP, "upload of solver factors failed");
    // zero every array once (fill kernels, not memset)
    for (int f = 0; f < QPAD_NFIELD && !rc; f++) rc = zero_dev(c, c->f[f], f1_size(c, c->ncomp[f]));
    double *hs[] = {c->h_psi, c->h_q_spe, c->h_q_beam};
    for (double *h : hs)
        if (!rc) rc = zero_dev(c, h, ns * f1_size(c, 1));
    double *hs3[] = {c->h_e, c->h_b, c->h_b_spe, c->h_e_spe, c->h_cu, c->h_b_beam};
    for (double *h : hs3)
        if (!rc) rc = zero_dev(c, h, ns * f1_size(c, 3));
    if (!rc) rc = zero_dev(c, c->conv_re, (size_t)g.nr + 2);
    if (!rc) rc = zero_dev(c, c->conv_im, (size_t)g.nr + 2);
    if (!rc) {
        memset(c->it_host, 0, sizeof(IterState));
        he = hipMemcpyAsync(c->it, c->it_host, sizeof(IterState), hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) rc = fail(QPAD_E_HIP, "iteration state upload failed");
    }
    if (!rc) rc = sync(c);
    if (rc) {
        qpad_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int qpad_destroy(qpad_ctx *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    seednet_destroy(c);
    for (int f = 0; f < QPAD_NFIELD; f++) dfree(c->f[f]);
    dfree(c->h_psi); dfree(c->h_e); dfree(c->h_b); dfree(c->h_b_spe); dfree(c->h_e_spe);
    dfree(c->h_cu); dfree(c->h_q_spe); dfree(c->h_q_beam); dfree(c->h_b_beam);
    dfree(c->conv_re); dfree(c->conv_im); dfree(c->conv_blk); dfree(c->it); dfree(c->fac); dfree(c->src);
    dfree(c->sc_ef);
    if (c->sc_ef_host) (void)hipHostFree(c->sc_ef_host);
    dfree(c->lz_ar); dfree(c->lz_ai); dfree(c->h_chi);
    if (c->it_host) (void)hipHostFree(c->it_host);
    if (c->it_report) (void)hipHostFree(c->it_report);
    if (c->dec_ev) (void)hipEventDestroy(c->dec_ev);
    for (int s = 0; s < c->nspecies; s++) species_free(c->sp[s]);
    for (int b = 0; b < c->nbeams; b++) beam_free(c->bm[b]);
    for (auto &p : c->pools)
        for (auto e : p.ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}


// Every check runs before the context is touched; the species becomes
// visible (nspecies grows) only once all of its buffers exist and its
// plasma is injected.  On any failure its buffers are freed and the context
// is left exactly as it was.
}  // extern "C"

static int species_create(qpad_ctx *c, const qpad_species_cfg *cfg, const qpad_neutral_cfg *nz, int32_t *sp_id);

extern "C" {

int qpad_species_new(qpad_ctx *c, const qpad_species_cfg *cfg, int32_t *sp_id)
{
    return species_create(c, cfg, nullptr, sp_id);
}

}  // extern "C"

// a plasma species (nz null) or a neutral gas (nz: its element, levels and
// n0)
static int species_create(qpad_ctx *c, const qpad_species_cfg *cfg, const qpad_neutral_cfg *nz, int32_t *sp_id)
{
    if (!c || !cfg) return fail(QPAD_E_INVALID, "null argument");
    if (c->nspecies >= kMaxSpecies) return fail(QPAD_E_INVALID, "too many species");
    if (cfg->ppc1 < 1 || cfg->ppc2 < 1 || cfg->num_theta < 1)
        return fail(QPAD_E_INVALID, "ppc and num_theta must be >= 1");
    if (cfg->push_type != QPAD_PUSH2_ROBUST && cfg->push_type != QPAD_PUSH2_STD && !is_pgc(cfg->push_type))
        return fail(QPAD_E_INVALID, "push_type must be QPAD_PUSH2_STD, QPAD_PUSH2_ROBUST, QPAD_PUSH2_STD_PGC or "
                                    "QPAD_PUSH2_ROBUST_PGC");
    if (cfg->mass == 0.0 || cfg->qm == 0.0) return fail(QPAD_E_INVALID, "q and m must be non-zero");
    if (c->subcyc && is_pgc(cfg->push_type))
        return fail(QPAD_E_STATE, "sub-cycling with the PGC pushers is not on the path");
    // npmax: fdist2d_class.f03:258-262 (xtra = 2)
    const int64_t nlat = (int64_t)c->g.nr * cfg->ppc1 * cfg->ppc2 * cfg->num_theta;
    if (nlat > 0x7fffffff / 2) return fail(QPAD_E_INVALID, "too many particles per species");
    size_t hist_ints = 0;
    TRY(sort_workspace(2 * nlat, c->g.nr, &hist_ints));
    const int s = c->nspecies;
    Species &sp = c->sp[s];
    memset((void *)&sp, 0, sizeof(Species));
    sp.cfg = *cfg;
    sp.qbm = cfg->qm / cfg->mass;
    sp.npmax = 2 * nlat;
    const size_t n = (size_t)sp.npmax;
    int rc = 0;
    double **dp[] = {&sp.x1, &sp.x2, &sp.p1, &sp.p2, &sp.p3, &sp.q, &sp.gamma, &sp.psi,
                     &sp.a_x1, &sp.a_x2, &sp.a_p1, &sp.a_p2, &sp.a_p3, &sp.a_q, &sp.a_gamma, &sp.a_psi};
    for (double **p : dp)
        if (!rc) rc = dalloc(p, n);
    if (!rc) rc = dalloc(&sp.id, n);
    if (!rc) rc = dalloc(&sp.a_id, n);
    if (!rc) rc = dalloc(&sp.key, n);
    if (!rc) rc = dalloc(&sp.key_alt, n);
    if (!rc) rc = dalloc(&sp.cell_start, (size_t)c->g.nr + 2);
    if (!rc) rc = dalloc(&sp.qn, f1_size(c, 1));
    if (!rc) rc = dalloc(&sp.P0, (size_t)c->g.nr * 8 * c->g.nm);
    if (!rc) rc = dalloc(&sp.P1, (size_t)c->g.nr * 8 * c->g.nm);
    sp.n_ov_max = (int)(sp.npmax / kDepChunk + 1);
    if (!rc) rc = dalloc(&sp.ov_start, (size_t)c->g.nr + 1);
    if (!rc) rc = dalloc(&sp.ov_desc, (size_t)sp.n_ov_max);
    if (!rc) rc = dalloc(&sp.O0, (size_t)sp.n_ov_max * 8 * c->g.nm);
    if (!rc) rc = dalloc(&sp.O1, (size_t)sp.n_ov_max * 8 * c->g.nm);
    if (!rc && (hipMemsetAsync(sp.ov_start, 0, sizeof(int32_t) * ((size_t)c->g.nr + 1), c->stream) != hipSuccess ||
                hipMemsetAsync(sp.ov_desc, 0, sizeof(int4) * (size_t)sp.n_ov_max, c->stream) != hipSuccess))
        rc = fail(QPAD_E_HIP, "hipMemsetAsync of the overflow rows failed");
    if (!rc) rc = dalloc(&sp.hist, hist_ints);
    if (!rc) rc = dalloc(&sp.tot, (size_t)c->g.nr + 1);
    if (!rc) rc = zero_dev(c, sp.qn, f1_size(c, 1));
    if (!rc && c->species_q) rc = species_q_alloc(c, s);
    if (!rc && (c->laser || is_pgc(cfg->push_type))) rc = species_chi_alloc(c, s);
    if (!rc && nz) rc = neutral_alloc(c, s, *nz);
    if (!rc) rc = species_restart(c, s);
    if (!rc) rc = sync(c);
    // a PGC species pushes in the laser envelope: a zero one until
    // qpad_set_laser (sim_lasers exists with nlasers = 0)
    if (!rc && is_pgc(cfg->push_type)) rc = laser_ensure(c);
    if (rc) {
        // keep the message of the first failure; wait for queued work on the
        // buffers before they are freed
        const std::string msg = g_err;
        (void)hipStreamSynchronize(c->stream);
        species_free(sp);
        memset((void *)&sp, 0, sizeof(Species));
        g_err = msg;
        return rc;
    }
    c->nspecies++;
    if (sp_id) *sp_id = s;
    return 0;
}

extern "C" {

int qpad_species_renew(qpad_ctx *c, int32_t s)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    TRY(species_restart(c, s));
    return sync(c);
}

// A neutral gas: a species of push type cfg->push_type with no particles
// until the field ionises it.  Checks first; on failure nothing changes.
int qpad_neutral_new(qpad_ctx *c, const qpad_neutral_cfg *cfg, int32_t *sp_id)
{
    if (!c || !cfg) return fail(QPAD_E_INVALID, "null argument");
    if (cfg->push_type != QPAD_PUSH2_STD && cfg->push_type != QPAD_PUSH2_ROBUST)
        return fail(QPAD_E_INVALID, "neutrals push with QPAD_PUSH2_STD or QPAD_PUSH2_ROBUST");
    if (cfg->ion_max < 1 || cfg->ion_max > 20 || !adk_row(cfg->element, cfg->ion_max))
        return fail(QPAD_E_INVALID, "neutral: element must be 1, 2, 3, 6, 7, 8, 10 or 18 and ion_max within "
                                    "its ionisation levels");
    if (!(cfg->n0 > 0.0)) return fail(QPAD_E_INVALID, "neutral: n0 must be > 0");
    if (1.0 * 1.0 * cfg->density < cfg->den_min)
        return fail(QPAD_E_INVALID, "neutral: density below den_min");
    qpad_species_cfg sc;
    memset(&sc, 0, sizeof(sc));
    sc.ppc1 = cfg->ppc1;
    sc.ppc2 = cfg->ppc2;
    sc.num_theta = cfg->num_theta;
    sc.qm = cfg->qm;
    sc.mass = cfg->mass;
    sc.density = cfg->density;
    sc.neutralized = 0;
    sc.den_min = cfg->den_min;
    sc.push_type = cfg->push_type;
    return species_create(c, &sc, cfg, sp_id);
}

}  // extern "C"

namespace qpad {
// the buffers of a beam of npmax particles; on failure everything allocated
// so far is released by the caller (beam_free)
static int beam_alloc(qpad_ctx *c, Beam &bm, int64_t npmax)
{
    bm.npmax = std::max<int64_t>(1, npmax);
    const size_t n = (size_t)bm.npmax;
    int rc = 0;
    double **dp[] = {&bm.x1, &bm.x2, &bm.x3, &bm.p1, &bm.p2, &bm.p3, &bm.q,
                     &bm.a_x1, &bm.a_x2, &bm.a_x3, &bm.a_p1, &bm.a_p2, &bm.a_p3, &bm.a_q};
    for (double **p : dp)
        if (!rc) rc = dalloc(p, n);
    if (!rc) rc = dalloc(&bm.key, n);
    if (!rc) rc = dalloc(&bm.key_alt, n);
    if (!rc) rc = dalloc(&bm.idx, n);
    if (!rc) rc = dalloc(&bm.idx_alt, n);
    if (!rc) rc = dalloc(&bm.cell_start, (size_t)c->g.nr * c->g.nxi + 2);
    if (!rc) {
        int bits = 1;
        while ((1u << bits) <= (uint32_t)(c->g.nr * c->g.nxi)) bits++;
        rc = sort_scratch(n, bits, &bm.sort_tmp_bytes);
        if (!rc) {
            hipError_t e = hipMalloc(&bm.sort_tmp, bm.sort_tmp_bytes);
            if (e != hipSuccess) rc = fail(QPAD_E_NOMEM, "beam sort scratch allocation failed");
        }
    }
    return rc;
}

// sort the injected beam by cell and publish it as beam b; on failure free
// its buffers and leave the context as it was
static int beam_finish(qpad_ctx *c, int b, int rc, int32_t *beam_id)
{
    Beam &bm = c->bm[b];
    if (!rc) rc = launch_beam_sort(c, b);
    if (!rc) rc = refresh_beam_npp(c, b);
    if (rc) {
        const std::string msg = g_err;
        (void)hipStreamSynchronize(c->stream);
        beam_free(bm);
        memset((void *)&bm, 0, sizeof(Beam));
        g_err = msg;
        return rc;
    }
    c->nbeams++;
    if (beam_id) *beam_id = b;
    return 0;
}
}  // namespace qpad

extern "C" {

int qpad_beam_new(qpad_ctx *c, const qpad_beam_cfg *cfg, int32_t *beam_id)
{
    if (!c || !cfg) return fail(QPAD_E_INVALID, "null argument");
    if (c->nbeams >= kMaxBeams) return fail(QPAD_E_INVALID, "too many beams");
    if (cfg->mass == 0.0) return fail(QPAD_E_INVALID, "beam mass must be non-zero");
    if (cfg->total_num < 0 || cfg->total_num > 0x3fffffff) return fail(QPAD_E_INVALID, "bad total_num");
    for (int d = 0; d < 3; d++)
        if (!(cfg->sigma[d] > 0.0) || !(cfg->range_hi[d] > cfg->range_lo[d]))
            return fail(QPAD_E_INVALID, "beam sigma and ranges must be positive");
    if ((int64_t)c->g.nr * c->g.nxi >= 0x7fffffff) return fail(QPAD_E_INVALID, "grid too large for beam cells");
    const int b = c->nbeams;
    Beam &bm = c->bm[b];
    memset((void *)&bm, 0, sizeof(Beam));
    bm.cfg = *cfg;
    bm.qbm = cfg->qm / cfg->mass;
    int rc = beam_alloc(c, bm, cfg->total_num * (cfg->quiet ? 2 : 1));
    if (!rc) rc = launch_beam_inject(c, b);
    return beam_finish(c, b, rc, beam_id);
}

int qpad_beam_new_arrays(qpad_ctx *c, double qm, double mass, int32_t evol, int64_t n, const double *x1,
                         const double *x2, const double *x3, const double *p1, const double *p2, const double *p3,
                         const double *q, int32_t *beam_id)
{
    if (!c) return fail(QPAD_E_INVALID, "null argument");
    if (c->nbeams >= kMaxBeams) return fail(QPAD_E_INVALID, "too many beams");
    if (mass == 0.0 || qm == 0.0) return fail(QPAD_E_INVALID, "beam q and m must be non-zero");
    if (n < 0 || n > 0x3fffffff) return fail(QPAD_E_INVALID, "bad beam particle count");
    if (n > 0 && (!x1 || !x2 || !x3 || !p1 || !p2 || !p3 || !q)) return fail(QPAD_E_INVALID, "null beam array");
    if ((int64_t)c->g.nr * c->g.nxi >= 0x7fffffff) return fail(QPAD_E_INVALID, "grid too large for beam cells");
    const int b = c->nbeams;
    Beam &bm = c->bm[b];
    memset((void *)&bm, 0, sizeof(Beam));
    bm.cfg.qm = qm;
    bm.cfg.mass = mass;
    bm.cfg.evol = evol;
    bm.qbm = qm / mass;
    int rc = beam_alloc(c, bm, n);
    const double *srcs[7] = {x1, x2, x3, p1, p2, p3, q};
    double *dsts[7] = {bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q};
    for (int f = 0; !rc && f < 7 && n; f++)
        if (hipMemcpyAsync(dsts[f], srcs[f], n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            rc = fail(QPAD_E_HIP, "copy of the beam particles failed");
    if (!rc) rc = sync(c);
    bm.npp = n;
    return beam_finish(c, b, rc, beam_id);
}

// the lattice beam, its density given per box point (tab) or by its profiles
static int beam_new_lattice(qpad_ctx *c, const qpad_beam_std_cfg *cfg, const double *den, int32_t k0, int32_t nk,
                            int32_t ni, int32_t *beam_id);

int qpad_beam_new_std(qpad_ctx *c, const qpad_beam_std_cfg *cfg, int32_t *beam_id)
{
    return beam_new_lattice(c, cfg, nullptr, 0, 0, 0, beam_id);
}

int qpad_beam_lattice_box(qpad_ctx *c, const qpad_beam_std_cfg *cfg, int32_t *k0, int32_t *nk, int32_t *ni)
{
    if (!c || !cfg || !k0 || !nk || !ni) return fail(QPAD_E_INVALID, "null argument");
    if (cfg->ppc[0] < 1 || cfg->ppc[1] < 1 || cfg->ppc[2] < 1 || cfg->num_theta < 1)
        return fail(QPAD_E_INVALID, "beam ppc and num_theta must be >= 1");
    int a = 0, b = 0, d = 0;
    TRY(lattice_range_box(c, *cfg, &a, &b, &d));
    *k0 = a;
    *nk = b;
    *ni = d;
    return 0;
}

int qpad_beam_new_std_density(qpad_ctx *c, const qpad_beam_std_cfg *cfg, const double *den, int32_t k0, int32_t nk,
                              int32_t ni, int32_t *beam_id)
{
    if (!c || !cfg || !den) return fail(QPAD_E_INVALID, "null argument");
    return beam_new_lattice(c, cfg, den, k0, nk, ni, beam_id);
}

static int beam_new_lattice(qpad_ctx *c, const qpad_beam_std_cfg *cfg, const double *den, int32_t k0, int32_t nk,
                            int32_t ni, int32_t *beam_id)
{
    if (!c || !cfg) return fail(QPAD_E_INVALID, "null argument");
    if (c->nbeams >= kMaxBeams) return fail(QPAD_E_INVALID, "too many beams");
    if (cfg->mass == 0.0 || cfg->qm == 0.0) return fail(QPAD_E_INVALID, "beam q and m must be non-zero");
    if (cfg->ppc[0] < 1 || cfg->ppc[1] < 1 || cfg->ppc[2] < 1 || cfg->num_theta < 1)
        return fail(QPAD_E_INVALID, "beam ppc and num_theta must be >= 1");
    if (cfg->geometry != QPAD_GEOM_CARTESIAN && cfg->geometry != QPAD_GEOM_CYLINDRICAL)
        return fail(QPAD_E_INVALID, "beam geometry must be cartesian or cylindrical");
    for (int d = 0; d < 3 && !den; d++) {
        const int pt = cfg->profile[d];
        if (pt != QPAD_PROF_UNIFORM && pt != QPAD_PROF_GAUSSIAN && pt != QPAD_PROF_PIECEWISE)
            return fail(QPAD_E_INVALID, "beam profile must be uniform, gaussian or piecewise-linear");
        if (pt == QPAD_PROF_GAUSSIAN && !(cfg->gauss_sigma[d] > 0.0))
            return fail(QPAD_E_INVALID, "gauss_sigma must be > 0");
        if (pt == QPAD_PROF_PIECEWISE) {
            if (cfg->pw_n[d] < 2 || cfg->pw_n[d] > QPAD_PW_MAX)
                return fail(QPAD_E_INVALID, "piecewise-linear profiles take 2 .. 32 points");
            for (int i = 1; i < cfg->pw_n[d]; i++)
                if (!(cfg->pw_x[d][i] > cfg->pw_x[d][i - 1]))
                    return fail(QPAD_E_INVALID, "Position array must be monotonically increasing!");
        }
        if (!(cfg->range_hi[d] >= cfg->range_lo[d])) return fail(QPAD_E_INVALID, "beam range upper < lower");
    }
    for (int d = 0; d < 2; d++) {
        if (cfg->offset_n[d] < 0 || cfg->offset_n[d] > QPAD_OFFSET_MAX)
            return fail(QPAD_E_INVALID, "perp_offset takes at most 8 coefficients");
        if (cfg->alpha[d] != 0.0 && !(cfg->twiss_beta[d] != 0.0))
            return fail(QPAD_E_INVALID, "Twiss alpha needs a non-zero beta");
    }
    if ((int64_t)c->g.nr * c->g.nxi >= 0x7fffffff) return fail(QPAD_E_INVALID, "grid too large for beam cells");
    const int b = c->nbeams;
    Beam &bm = c->bm[b];
    memset((void *)&bm, 0, sizeof(Beam));
    // the fields the push and the host mirror read (charge sign, mass, evolution)
    bm.cfg.qm = cfg->qm;
    bm.cfg.mass = cfg->mass;
    bm.cfg.gamma = cfg->gamma;
    bm.cfg.quiet = cfg->quiet;
    bm.cfg.evol = cfg->evol;
    bm.cfg.seed = cfg->seed;
    bm.qbm = cfg->qm / cfg->mass;
    LatticePlan pl;
    memset((void *)&pl, 0, sizeof(pl));
    LatticeTable tab{nullptr, k0, nk, ni};
    double *dtab = nullptr;
    int rc = 0;
    if (den) {
        // the table's box must hold every lattice point inside the ranges
        int r0 = 0, rn = 0, ri = 0;
        rc = lattice_range_box(c, *cfg, &r0, &rn, &ri);
        if (!rc && (k0 < 1 || nk < 0 || ni < 0 || k0 + nk - 1 > c->g.nxi || ni > c->g.nr ||
                    (rn > 0 && (k0 > r0 || k0 + nk < r0 + rn || ni < ri))))
            rc = fail(QPAD_E_INVALID, "lattice density box: xi cells k0 .. k0+nk-1 within 1 .. nxi and r cells "
                                      "1 .. ni <= nr, holding the box of the ranges (qpad_beam_lattice_box)");
        const int64_t n = (int64_t)nk * cfg->num_theta * ni * cfg->ppc[0] * cfg->ppc[1] * cfg->ppc[2];
        if (!rc && n > 0) {
            rc = dalloc(&dtab, (size_t)n);
            if (!rc && hipMemcpyAsync(dtab, den, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                           hipSuccess)
                rc = fail(QPAD_E_HIP, "copy of the lattice density failed");
        }
        tab.den = dtab;
    }
    if (!rc) rc = lattice_count(c, *cfg, &pl, den ? &tab : nullptr);
    if (!rc && pl.kept > 0x3fffffff) rc = fail(QPAD_E_INVALID, "lattice beam: too many particles");
    if (!rc) rc = beam_alloc(c, bm, pl.kept * (cfg->quiet ? 2 : 1));
    if (!rc) rc = lattice_write(c, *cfg, pl, bm, den ? &tab : nullptr);
    if (!rc) rc = sync(c);
    (void)hipStreamSynchronize(c->stream);
    dfree(dtab);
    lattice_free(&pl);
    return beam_finish(c, b, rc, beam_id);
}

// ---- fused step ----

}  // extern "C"

namespace qpad {
// the 3D-step head without the final host synchronisation
static int step_begin_enqueue(qpad_ctx *c)
{
    TRY(seednet_quiesce(c));
    // q_beam%as(0); beam3d%qdp (simulation_class.f03:297-306)
    for (int b = 0; b < c->nbeams; b++) TRY(launch_beam_sort(c, b));
    if (c->nbeams > 0) TRY(launch_beam_qdeposit(c));
    else TRY(zero_dev(c, c->h_q_beam, ((size_t)c->g.nxi + 2) * f1_size(c, 1)));
    // beam B_perp for every slice (hoisted out of the slice loop: it depends
    // on the q_beam history only)
    TRY(launch_bt_beam_all(c));
    TRY(zero_dev(c, c->h_q_spe, ((size_t)c->g.nxi + 2) * f1_size(c, 1)));
    const int zf[] = {QPAD_F_B, QPAD_F_E, QPAD_F_B_SPE, QPAD_F_E_SPE, QPAD_F_PSI, QPAD_F_CU, QPAD_F_ACU,
                      QPAD_F_AMU, QPAD_F_Q_SPE};
    for (int f : zf) TRY(zero_dev(c, c->f[f], f1_size(c, c->ncomp[f])));
    TRY(seednet_step_begin(c));
    c->spec_guess = 2;
    c->step_open = true;
    return 0;
}
}  // namespace qpad

extern "C" {

int qpad_step3d_begin(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    TRY(step_begin_enqueue(c));
    return sync(c);
}

}  // extern "C"

namespace qpad {

// One xi-slice (simulation_class.f03:342-469), enqueued without host
// synchronisation: all iter_max predictor-corrector iterations are queued and
// each of their kernels returns at once after the device-side convergence
// test has set IterState::done.
// overlap: also queue the next slice's seed behind this slice's epilogue
// (qpad_step3d, where nothing can change the fields between slices)
// A slice's queuing as a resumable state machine, so that the blocking form
// (slice_enqueue) and the batch's event-driven form queue the same kernels.
struct SliceRun {
    int j = 0;            // slice being queued
    int queued = 0;       // iterations queued so far
    int upto = 0;         // iterations to queue before the next decision
    int known_iters = 0;  // iterations the slice needed, once known
    bool waiting = false; // a deciding deposit is in flight (dec_ev)
    bool overlap = false; // network overlap (qpad_step3d only)
    int sub = 1, nsub = 1; // sub-cycle being queued, sub-cycles of the slice
};

static int slice_finish(qpad_ctx *c, SliceRun &st);

// Queue the iterations up to st.upto, then (unless iter_max is reached) the
// next iteration, whose deposit decides on the one before it and reports to
// the host behind dec_ev; or finish the slice.
static int slice_advance(qpad_ctx *c, SliceRun &st)
{
    const int j = st.j, imax = c->g.iter_max;
    for (int l = st.queued + 1; l <= st.upto; l++) {
        const bool first = l == 1;
        TRY(launch_amjdeposit_all(c, l, true, first));
        hipEvent_t stop = nullptr;
        if (first) time_begin(c, 2, &stop);
        TRY(launch_iteration_fields(c, j, l));
        if (first) time_end(c, 2, stop);
    }
    st.queued = std::max(st.queued, st.upto);
    if (st.queued >= imax) return slice_finish(c, st);
    // iteration queued + 1: its deposit decides on iteration `queued`
    const int l = st.queued + 1;
    TRY(launch_amjdeposit_all(c, l, true, false, true));
    QPAD_HIP_CHECK(hipEventRecord(c->dec_ev, c->stream));
    TRY(launch_iteration_fields(c, j, l));
    st.queued = l;
    st.waiting = true;
    return 0;
}

// the deciding deposit has completed: finish the slice, or queue on
static int slice_decided(qpad_ctx *c, SliceRun &st)
{
    st.waiting = false;
    if (c->it_report->done) {
        st.known_iters = c->it_report->iters;
        return slice_finish(c, st);
    }
    st.upto = st.queued;
    return slice_advance(c, st);
}

// rho - Jz of every species, psi, Bz from the J predictor, 'record', then
// the first iterations.  Each iteration's kernels return at once once the
// device-side test has set IterState::done, so queuing more iterations than
// needed changes no result, only costs launches.  Adaptive mode queues as
// many iterations as the previous slice needed (plus the margin); after
// that, the deposit of each further iteration (which decides whether the
// one before it converged) reports the decision to the host, with that
// iteration's field kernels already queued behind it, and the host stops
// queuing at the first converged report.  Fixed mode queues all iter_max
// iterations.  Timing events (when enabled) bracket only the first
// iteration, which always does work: an event per launch would cost a
// quarter of the step.  With sub-cycling this is the start of every
// sub-cycle (simulation_subcyc_class.f03:282-300).
static int subcycle_start(qpad_ctx *c, SliceRun &st)
{
    const int j = st.j;
    const bool net = seednet_active(c);
    const double *seed = nullptr;
    TRY(launch_qdeposit_all(c));
    if (c->species_q)
        for (int s = 0; s < c->nspecies; s++) TRY(launch_species_q_start(c, s));
    // neutral%ion_deposit (simulation_class.f03:351-354): rho_ion, which
    // the sources add to q_spe after the electrons' charge
    for (int s = 0; s < c->nspecies; s++)
        if (c->sp[s].neutral) TRY(launch_ion_deposit(c, s));
    TRY(launch_slice_sources(c));
    TRY(launch_start_solves(c));
    // species2d%interp_psi of the standard-pusher species
    // (simulation_class.f03:357-359); QPAD's loop leaves out the neutrals,
    // whose electrons keep the psi = 0 of their release
    for (int s = 0; s < c->nspecies; s++)
        if (!c->sp[s].neutral) TRY(launch_interp_psi(c, s));
    // the first kernel that reads B_perp waits for the seed
    if (net) TRY(seednet_before_sources(c, j, &seed));
    TRY(launch_start_eperp(c, j, seed));
    const int imax = c->g.iter_max;
    st.queued = 0;
    st.waiting = false;
    st.known_iters = imax;
    // (at most imax - 1 up front, so that the deposit of iteration imax still
    // reports and the next slice's guess keeps following the iterations)
    st.upto = c->adaptive ? std::max(1, std::min(imax - 1, c->spec_guess + c->spec_margin)) : imax;
    return slice_advance(c, st);
}

// the sub-cycles of slice j (get_exp_fac_max and get_subcyc_step,
// simulation_subcyc_class.f03:270-276): the 2D step and the clamp of its
// deposits and pushes
static int subcycle_plan(qpad_ctx *c, SliceRun &st)
{
    double ef = 1.0, dsub = c->g.dxi;
    TRY(launch_exp_fac_max(c, &ef));
    if (!(ef < INFINITY)) return fail(QPAD_E_STATE, "sub-cycling: a particle has gamma - p_z = 0");
    const int n = subcyc_step(ef, c->sc_fac_max, c->g.dxi, c->sc_dt_min, &dsub);
    if (n < 1) return fail(QPAD_E_STATE, "sub-cycling: no sub-cycle fits dt_2d_min");
    st.sub = 1;
    st.nsub = n;
    c->dt2d = dsub;
    c->clamp2d = c->sc_fac_clamped;
    c->last_nsub = n;
    c->last_dsub = dsub;
    c->last_ef = ef;
    return launch_iter_fold(c, true);
}

static int slice_start(qpad_ctx *c, int j, bool overlap, SliceRun &st)
{
    // the learned seed (when on) is computed on the network's stream from
    // the fields left by slice j - 1, beside the previous slice's push and
    // sort and this slice's charge deposit, sources and solves
    if (seednet_active(c) && c->net_pre != j) TRY(seednet_after_epilogue(c, 0, j));
    st = SliceRun();
    st.j = j;
    st.overlap = overlap;
    if (c->subcyc) TRY(subcycle_plan(c, st));
    // laser_all of slice j (simulation_class.f03:362-367)
    if (c->laser) TRY(launch_laser_slice(c, j));
    return subcycle_start(c, st);
}

// slice epilogue: cbq, e_spe, history (simulation_class.f03:401-456), then
// the push and the sort.  The network's work behind the epilogue (the next
// seed, training) is queued after the push and the sort: the host reaches
// this point only once the slice's convergence is known, and queuing the
// network first left the main stream idle while the host was busy with it.
// A sub-cycle before the slice's last only pushes (the neutrals' update,
// push_u, clamp_exp_fac, push_x: simulation_subcyc_class.f03:337-364) and
// starts the next; the last one's epilogue comes before its push, which
// changes nothing it reads (QPAD closes the slice after the push).
static int slice_finish(qpad_ctx *c, SliceRun &st)
{
    const int j = st.j;
    if (c->adaptive) c->spec_guess = st.known_iters;
    if (st.sub < st.nsub) {
        TRY(launch_iter_fold(c, false));
        for (int s = 0; s < c->nspecies; s++)
            if (c->sp[s].neutral) TRY(launch_neutral_update(c, s));
        TRY(launch_push_all(c));
        TRY(launch_sort_all(c));
        st.sub++;
        return subcycle_start(c, st);
    }
    TRY(launch_epilogue(c, j));
    if (c->species_q)
        for (int s = 0; s < c->nspecies; s++) TRY(launch_species_cbq(c, s, j));
    // sim_lasers%deposit_chi (simulation_class.f03:400-401): the particles'
    // psi of the last current deposit, before the push
    if (c->laser) TRY(launch_deposit_chi(c, j));
    // neutral%cbq's rho_ion copy, then neutral%update (simulation_class.f03:
    // 406-407, 444-445): the levels from the final E and the electrons they
    // release, which the push below moves with the rest
    for (int s = 0; s < c->nspecies; s++) {
        Species &sp = c->sp[s];
        if (!sp.neutral) continue;
        const size_t n1 = f1_size(c, 1);
        TRY(launch_copy(c, sp.h_rho + (size_t)j * n1, sp.qn, n1));
        TRY(launch_neutral_update(c, s));
    }
    const bool net = seednet_active(c);
    if (net) TRY(seednet_mark(c));
    TRY(launch_push_all(c));
    TRY(launch_sort_all(c));
    if (net) TRY(seednet_after_epilogue(c, j, (st.overlap && j < c->g.nxi) ? j + 1 : 0, true));
    c->dt2d = c->g.dxi;
    c->clamp2d = 0.0;
    st.j = 0;
    return 0;
}

static int slice_enqueue(qpad_ctx *c, int j, bool overlap)
{
    SliceRun st;
    TRY(slice_start(c, j, overlap, st));
    while (st.waiting) {
        QPAD_HIP_CHECK(hipEventSynchronize(c->dec_ev));
        TRY(slice_decided(c, st));
    }
    return 0;
}

static int read_iter_state(qpad_ctx *c)
{
    QPAD_HIP_CHECK(hipMemcpyAsync(c->it_host, c->it, sizeof(IterState), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

}  // namespace qpad

extern "C" {

int qpad_slice(qpad_ctx *c, int32_t j, int32_t *iters_out)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (j < 1 || j > c->g.nxi) return fail(QPAD_E_INVALID, "slice index out of range");
    TRY(slice_enqueue(c, j, false));
    TRY(read_iter_state(c));
    if (iters_out) *iters_out = c->it_host->iters + (c->subcyc ? c->it_host->sub_iters : 0);
    return 0;
}

int qpad_step3d_end(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    for (int b = 0; b < c->nbeams; b++) TRY(launch_beam_push(c, b));
    for (int s = 0; s < c->nspecies; s++) TRY(species_restart(c, s));
    c->step_open = false;
    return sync(c);
}

int qpad_step_iterations(qpad_ctx *c, int32_t *total_iters, int32_t *slices_at_iter_max)
{
    if (!c || !total_iters || !slices_at_iter_max) return fail(QPAD_E_INVALID, "null argument");
    *total_iters = c->it_host->total_iters;
    *slices_at_iter_max = c->it_host->at_max;
    return 0;
}

int qpad_step3d(qpad_ctx *c, int32_t *total_iters_out)
{
    TRY(qpad_step3d_begin(c));
    TRY(read_iter_state(c));
    c->it_host->total_iters = 0;
    c->it_host->at_max = 0;
    QPAD_HIP_CHECK(hipMemcpyAsync(c->it, c->it_host, sizeof(IterState), hipMemcpyHostToDevice, c->stream));
    for (int j = 1; j <= c->g.nxi; j++) TRY(slice_enqueue(c, j, true));
    TRY(read_iter_state(c));
    const int32_t tot = c->it_host->total_iters;
    TRY(qpad_step3d_end(c));
    if (total_iters_out) *total_iters_out = tot;
    return 0;
}

int qpad_get_launch_shape(qpad_ctx *c, qpad_launch_shape *out)
{
    if (!c || !out) return fail(QPAD_E_INVALID, "null argument");
    out->dep_cells = c->dep_cells;
    out->dep_mfma = c->dep_mfma;
    return 0;
}

int qpad_set_launch_shape(qpad_ctx *c, const qpad_launch_shape *shape)
{
    if (!c || !shape) return fail(QPAD_E_INVALID, "null argument");
    if (shape->dep_cells != 2 && shape->dep_cells != 4 && shape->dep_cells != 8)
        return fail(QPAD_E_INVALID, "dep_cells must be 2, 4 or 8");
    if (shape->dep_mfma != 0 && shape->dep_mfma != 1) return fail(QPAD_E_INVALID, "dep_mfma must be 0 or 1");
    c->dep_cells = shape->dep_cells;
    c->dep_mfma = shape->dep_mfma;
    return 0;
}

// Every candidate shape runs the current deposit of all species (first-
// iteration instance, no gamma/psi stores) and their charge deposit, the
// same launches the slice queues, on the current particles and fields,
// `reps` timed launches each after one untimed.  Both
// kernels only write the per-cell partials, which every slice recomputes
// before use, so tuning changes no result.
int qpad_tune_launch(qpad_ctx *c, int32_t reps, double *ms_out)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (reps < 1 || reps > 1000) return fail(QPAD_E_INVALID, "reps must be 1..1000");
    TRY(seednet_quiesce(c));
    TRY(sync(c));
    const int cand[3] = {2, 4, 8};
    const int saved = c->dep_cells;
    double best = 0.0;
    int best_c = saved;
    hipEvent_t e0, e1;
    QPAD_HIP_CHECK(hipEventCreate(&e0));
    QPAD_HIP_CHECK(hipEventCreate(&e1));
    int rc = 0;
    for (int k = 0; k < 3 && !rc; k++) {
        c->dep_cells = cand[k];
        for (int r = 0; r <= reps && !rc; r++) {
            if (r == 1 && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(QPAD_E_HIP, "hipEventRecord");
            // the all-species launches slice_advance and slice_start queue
            rc = launch_amjdeposit_all(c, 1, false, false);
            if (!rc) rc = launch_qdeposit_all(c);
        }
        float ms = 0.0f;
        if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(QPAD_E_HIP, "tuning events failed");
        const double per = (double)ms / reps;
        if (ms_out) ms_out[k] = per;
        if (!rc && (k == 0 || per < best)) {
            best = per;
            best_c = cand[k];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->dep_cells = rc ? saved : best_c;
    return rc;
}

int qpad_iteration_policy(qpad_ctx *c, int32_t adaptive, int32_t margin)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (margin < -c->g.iter_max || margin > c->g.iter_max) return fail(QPAD_E_INVALID, "margin out of range");
    c->adaptive = adaptive != 0;
    c->spec_margin = margin;
    return 0;
}

// ---- fine-grained drop-in entry points ----

int qpad_species_qdp(qpad_ctx *c, int32_t s)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    TRY(launch_qdeposit(c, s));
    if (c->species_q) TRY(launch_species_q_start(c, s));
    TRY(launch_combine_q(c, s, c->f[QPAD_F_Q_SPE], 1.0, true));
    if (c->sp[s].cfg.neutralized) TRY(launch_add_qn_into_qspe(c, s));
    return sync(c);
}

static int copy_rows(qpad_ctx *c, double *dev, int ncomp, double *host_out, const double *host_in);

int qpad_species_cbq(qpad_ctx *c, int32_t s, int32_t pos)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    if (!c->species_q) return fail(QPAD_E_STATE, "per-species charge is off (qpad_species_charge_history)");
    if (pos < 1 || pos > c->g.nxi) return fail(QPAD_E_INVALID, "slice index out of range");
    TRY(launch_species_cbq(c, s, pos));
    if (c->sp[s].neutral) {   // neutral%cbq also keeps rho_ion (neutral_class.f03:1158-1176)
        const size_t n1 = f1_size(c, 1);
        TRY(launch_copy(c, c->sp[s].h_rho + (size_t)pos * n1, c->sp[s].qn, n1));
    }
    return sync(c);
}

int qpad_species_charge_history(qpad_ctx *c, int32_t enable)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (enable && !c->species_q) {
        for (int s = 0; s < c->nspecies; s++) {
            const int rc = species_q_alloc(c, s);
            if (rc) {
                for (int t = 0; t <= s; t++) {
                    dfree(c->sp[t].qsl);
                    dfree(c->sp[t].h_q);
                    c->sp[t].qsl = c->sp[t].h_q = nullptr;
                }
                return rc;
            }
        }
    }
    c->species_q = enable != 0;
    return sync(c);
}

int qpad_get_species_charge(qpad_ctx *c, int32_t s, int32_t slice, double *out)
{
    if (!c || s < 0 || s >= c->nspecies || !out) return fail(QPAD_E_INVALID, "bad species");
    if (!c->sp[s].h_q) return fail(QPAD_E_STATE, "per-species charge is off (qpad_species_charge_history)");
    if (slice < 0 || slice > c->g.nxi + 1) return fail(QPAD_E_INVALID, "bad history request");
    return copy_rows(c, c->sp[s].h_q + (size_t)slice * f1_size(c, 1), 1, out, nullptr);
}

int qpad_species_amjdp(qpad_ctx *c, int32_t s, double dxi)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    if (!(dxi > 0.0 && dxi < INFINITY)) return fail(QPAD_E_INVALID, "dxi must be > 0");
    c->dt2d = dxi;
    const int rc = launch_amjdeposit(c, s, 0, true, false);
    c->dt2d = c->g.dxi;
    TRY(rc);
    // add this species' cu/acu/amu into the totals
    TRY(launch_combine_j_species(c, s));
    return sync(c);
}

int qpad_set_laser(qpad_ctx *c, const double *ar_hist, const double *ai_hist)
{
    if (!c || !ar_hist || !ai_hist) return fail(QPAD_E_INVALID, "null argument");
    if (c->subcyc) return fail(QPAD_E_STATE, "sub-cycling with a laser is not on the path");
    TRY(laser_ensure(c));
    // packed host [(nxi + 3)][nm][nr + 2] -> padded device rows of length ld
    const size_t rows = ((size_t)c->g.nxi + 3) * c->g.nm, w = (size_t)c->g.nr + 2;
    QPAD_HIP_CHECK(hipMemcpy2DAsync(c->lz_ar, c->g.ld * sizeof(double), ar_hist, w * sizeof(double),
                                    w * sizeof(double), rows, hipMemcpyHostToDevice, c->stream));
    QPAD_HIP_CHECK(hipMemcpy2DAsync(c->lz_ai, c->g.ld * sizeof(double), ai_hist, w * sizeof(double),
                                    w * sizeof(double), rows, hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

int qpad_laser_slice(qpad_ctx *c, int32_t j)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (j < 1 || j > c->g.nxi) return fail(QPAD_E_INVALID, "slice index out of range");
    if (!c->laser) return fail(QPAD_E_STATE, "no laser: call qpad_set_laser or add a PGC species first");
    TRY(launch_laser_slice(c, j));
    return sync(c);
}

int qpad_deposit_chi(qpad_ctx *c, int32_t j)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (j < 0 || j > c->g.nxi + 1) return fail(QPAD_E_INVALID, "slice index out of range");
    if (!c->laser) return fail(QPAD_E_STATE, "no laser: call qpad_set_laser or add a PGC species first");
    TRY(launch_deposit_chi(c, j));
    return sync(c);
}

int qpad_neutral_ion_deposit(qpad_ctx *c, int32_t s)
{
    if (!c || s < 0 || s >= c->nspecies || !c->sp[s].neutral) return fail(QPAD_E_INVALID, "bad neutral");
    TRY(launch_ion_deposit(c, s));
    TRY(launch_add_qn_into_qspe(c, s));
    return sync(c);
}

int qpad_neutral_update(qpad_ctx *c, int32_t s)
{
    if (!c || s < 0 || s >= c->nspecies || !c->sp[s].neutral) return fail(QPAD_E_INVALID, "bad neutral");
    return launch_neutral_update(c, s);
}

int qpad_get_neutral_levels(qpad_ctx *c, int32_t s, double *out)
{
    if (!c || s < 0 || s >= c->nspecies || !c->sp[s].neutral || !out) return fail(QPAD_E_INVALID, "bad neutral");
    const Species &sp = c->sp[s];
    const size_t n = (size_t)sp.cfg.num_theta * c->g.nr * (sp.multi_max + 2);
    QPAD_HIP_CHECK(hipMemcpyAsync(out, sp.levels, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int qpad_get_ion_density(qpad_ctx *c, int32_t s, int32_t slice, double *out)
{
    if (!c || s < 0 || s >= c->nspecies || !c->sp[s].neutral || !out) return fail(QPAD_E_INVALID, "bad neutral");
    if (slice < 0 || slice > c->g.nxi + 1) return fail(QPAD_E_INVALID, "slice index out of range");
    return copy_rows(c, c->sp[s].h_rho + (size_t)slice * f1_size(c, 1), 1, out, nullptr);
}

int qpad_neutral_adk(int32_t element, int32_t level, double *out)
{
    const double *r = adk_row(element, level);
    if (!r || !out) return fail(QPAD_E_INVALID, "no ADK coefficients for this element and level");
    for (int t = 0; t < 3; t++) out[t] = r[t];
    return 0;
}

int qpad_species_interp_psi(qpad_ctx *c, int32_t s)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    if (c->sp[s].neutral) return fail(QPAD_E_INVALID, "a neutral has no interp_psi (its electrons keep psi = 0)");
    TRY(launch_interp_psi(c, s));
    return sync(c);
}

// the push of species s with 2D step dxi and expansion-factor clamp (0: none)
static int species_push(qpad_ctx *c, int32_t s, double dxi, double clamp)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    if (!(dxi > 0.0 && dxi < INFINITY)) return fail(QPAD_E_INVALID, "dxi must be > 0");
    c->dt2d = dxi;
    c->clamp2d = clamp;
    int rc = launch_push(c, s);
    c->dt2d = c->g.dxi;
    c->clamp2d = 0.0;
    TRY(rc);
    TRY(launch_sort(c, s));
    return sync(c);
}

int qpad_species_push(qpad_ctx *c, int32_t s, double dxi) { return species_push(c, s, dxi, 0.0); }

int qpad_species_push_clamped(qpad_ctx *c, int32_t s, double dxi, double exp_fac_clamped)
{
    if (!(exp_fac_clamped > 1.0 && exp_fac_clamped < INFINITY))
        return fail(QPAD_E_INVALID, "expansion_fac_clamped must be > 1");
    if (c && s >= 0 && s < c->nspecies && is_pgc(c->sp[s].cfg.push_type))
        return fail(QPAD_E_INVALID, "the clamp is for the standard and robust pushers");
    return species_push(c, s, dxi, exp_fac_clamped);
}

int qpad_set_subcycling(qpad_ctx *c, int32_t on, double exp_fac_max, double exp_fac_clamped, double dt_2d_min)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    if (on != 0 && on != 1) return fail(QPAD_E_INVALID, "on must be 0 or 1");
    if (!on) {
        c->subcyc = false;
        return 0;
    }
    if (!(exp_fac_max > 0.0 && exp_fac_max < INFINITY)) return fail(QPAD_E_INVALID, "expansion_fac_max must be > 0");
    if (!(exp_fac_clamped > 1.0 && exp_fac_clamped < INFINITY))
        return fail(QPAD_E_INVALID, "expansion_fac_clamped must be > 1");
    if (!(dt_2d_min > 0.0 && dt_2d_min <= c->g.dxi && c->g.dxi / dt_2d_min <= 4096.0))
        return fail(QPAD_E_INVALID, "dt_2d_min must be in [dxi / 4096, dxi]");
    if (c->laser) return fail(QPAD_E_STATE, "sub-cycling with a laser is not on the path");
    for (int s = 0; s < c->nspecies; s++)
        if (is_pgc(c->sp[s].cfg.push_type))
            return fail(QPAD_E_STATE, "sub-cycling with the PGC pushers is not on the path");
    if (c->net) return fail(QPAD_E_STATE, "sub-cycling with the seed network is not on the path");
    c->subcyc = true;
    c->sc_fac_max = exp_fac_max;
    c->sc_fac_clamped = exp_fac_clamped;
    c->sc_dt_min = dt_2d_min;
    return 0;
}

int qpad_species_set_profile(qpad_ctx *c, int32_t s, const double *den_perp, double den_lon)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    Species &sp = c->sp[s];
    if (sp.neutral) return fail(QPAD_E_INVALID, "neutral profiles are uniform on this path");
    if (!(den_lon >= 0.0 && den_lon < INFINITY)) return fail(QPAD_E_INVALID, "den_lon must be finite and >= 0");
    const int64_t n = (int64_t)sp.cfg.num_theta * c->g.nr * sp.cfg.ppc1 * sp.cfg.ppc2;
    if (!den_perp) {
        // uniform again: every lattice point, no longitudinal factor
        (void)hipStreamSynchronize(c->stream);
        dfree(sp.den_perp); dfree(sp.inj_keep); dfree(sp.inj_off);
        if (sp.inj_tmp) (void)hipFree(sp.inj_tmp);
        if (sp.inj_host) (void)hipHostFree(sp.inj_host);
        sp.den_perp = nullptr;
        sp.inj_keep = sp.inj_off = nullptr;
        sp.inj_tmp = nullptr;
        sp.inj_host = nullptr;
        sp.den_lon = 1.0;
        return 0;
    }
    for (int64_t l = 0; l < n; l++)
        if (!(den_perp[l] < INFINITY && den_perp[l] > -INFINITY))
            return fail(QPAD_E_INVALID, "den_perp must be finite");
    if (!sp.den_perp) {
        int rc = dalloc(&sp.den_perp, (size_t)n);
        if (!rc) rc = dalloc(&sp.inj_keep, (size_t)n + 1);
        if (!rc) rc = dalloc(&sp.inj_off, (size_t)n + 1);
        if (!rc) rc = inject_scan_bytes(n, &sp.inj_tmp_bytes);
        if (!rc && hipMalloc(&sp.inj_tmp, std::max<size_t>(sp.inj_tmp_bytes, 1)) != hipSuccess)
            rc = fail(QPAD_E_NOMEM, "hipMalloc of the injection scan scratch failed");
        if (!rc && hipHostMalloc((void **)&sp.inj_host, sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
            rc = fail(QPAD_E_NOMEM, "hipHostMalloc failed");
        if (rc) {
            dfree(sp.den_perp); dfree(sp.inj_keep); dfree(sp.inj_off);
            if (sp.inj_tmp) (void)hipFree(sp.inj_tmp);
            sp.den_perp = nullptr;
            sp.inj_keep = sp.inj_off = nullptr;
            sp.inj_tmp = nullptr;
            return rc;
        }
    }
    QPAD_HIP_CHECK(hipMemcpyAsync(sp.den_perp, den_perp, (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
    sp.den_lon = den_lon;
    return sync(c);
}

int qpad_exp_fac_max(qpad_ctx *c, double *out)
{
    if (!c || !out) return fail(QPAD_E_INVALID, "null argument");
    return launch_exp_fac_max(c, out);
}

int qpad_last_subcycles(qpad_ctx *c, int32_t *n, double *dxi_sub, double *exp_fac)
{
    if (!c || !n || !dxi_sub || !exp_fac) return fail(QPAD_E_INVALID, "null argument");
    *n = c->last_nsub;
    *dxi_sub = c->last_dsub;
    *exp_fac = c->last_ef;
    return 0;
}

int qpad_field_solve_psi(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    TRY(launch_psi_from_qspe(c));
    return sync(c);
}

static bool valid_fid(int f) { return f >= 0 && f < QPAD_NFIELD; }

int qpad_field_solve_ez(qpad_ctx *c, int32_t dst, int32_t jay)
{
    if (!c || !valid_fid(dst) || !valid_fid(jay) || c->ncomp[dst] != 3 || c->ncomp[jay] != 3)
        return fail(QPAD_E_INVALID, "bad field ids");
    TRY(launch_ez_only(c, dst, jay));
    return sync(c);
}

int qpad_field_solve_et(qpad_ctx *c, int32_t dst, int32_t b)
{
    if (!c || !valid_fid(dst) || !valid_fid(b) || c->ncomp[dst] != 3 || c->ncomp[b] != 3)
        return fail(QPAD_E_INVALID, "bad field ids");
    TRY(launch_et_only(c, dst, b));
    return sync(c);
}

int qpad_field_solve_bz(qpad_ctx *c, int32_t dst, int32_t jay)
{
    if (!c || !valid_fid(dst) || !valid_fid(jay) || c->ncomp[dst] != 3 || c->ncomp[jay] != 3)
        return fail(QPAD_E_INVALID, "bad field ids");
    TRY(launch_bz_only(c, dst, jay));
    return sync(c);
}

int qpad_field_solve_bt_beam(qpad_ctx *c, int32_t j)
{
    if (!c || j < 1 || j > c->g.nxi) return fail(QPAD_E_INVALID, "bad slice");
    TRY(launch_bt_beam_one(c, j));
    return sync(c);
}

int qpad_field_solve_djdxi(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    TRY(launch_djdxi_only(c));
    return sync(c);
}

int qpad_field_solve_bt_iter(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    TRY(launch_bperp_bz(c, true, false, false, false));
    return sync(c);
}

int qpad_field_add_b(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    TRY(launch_add_b(c, 0));
    return sync(c);
}

int qpad_convergence_tester(qpad_ctx *c, int32_t op, double *rel, double *ab)
{
    if (!c || (op != 0 && op != 1)) return fail(QPAD_E_INVALID, "bad convergence_tester operation");
    if (op == 0) {
        TRY(launch_psi_bz(c, false, false, true));
        return sync(c);
    }
    TRY(launch_compare_only(c));
    QPAD_HIP_CHECK(hipMemcpyAsync(c->it_host, c->it, sizeof(IterState), hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    if (rel) *rel = c->it_host->rel;
    if (ab) *ab = c->it_host->abs_;
    return 0;
}

int qpad_beam_qdp(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    for (int b = 0; b < c->nbeams; b++) TRY(launch_beam_sort(c, b));
    if (c->nbeams > 0) TRY(launch_beam_qdeposit(c));
    return sync(c);
}

int qpad_beam_push(qpad_ctx *c)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    for (int b = 0; b < c->nbeams; b++) TRY(launch_beam_push(c, b));
    return sync(c);
}

// ---- state exchange ----

static double *hist_ptr(qpad_ctx *c, int fid, int *ncomp)
{
    switch (fid) {
    case QPAD_F_PSI: *ncomp = 1; return c->h_psi;
    case QPAD_F_E: *ncomp = 3; return c->h_e;
    case QPAD_F_B: *ncomp = 3; return c->h_b;
    case QPAD_F_B_SPE: *ncomp = 3; return c->h_b_spe;
    case QPAD_F_E_SPE: *ncomp = 3; return c->h_e_spe;
    case QPAD_F_CU: *ncomp = 3; return c->h_cu;
    case QPAD_F_Q_SPE: *ncomp = 1; return c->h_q_spe;
    case QPAD_F_Q_BEAM: *ncomp = 1; return c->h_q_beam;
    case QPAD_F_B_BEAM: *ncomp = 3; return c->h_b_beam;
    case QPAD_F_CHI: *ncomp = 1; return c->h_chi;   // null before the laser state exists
    default: return nullptr;
    }
}

// packed host layout [ncomp][nm][nr+2] <-> padded device rows of length ld
static int copy_rows(qpad_ctx *c, double *dev, int ncomp, double *host_out, const double *host_in)
{
    const size_t rows = (size_t)ncomp * c->g.nm, w = (size_t)c->g.nr + 2;
    if (host_out) {
        QPAD_HIP_CHECK(hipMemcpy2DAsync(host_out, w * sizeof(double), dev, c->g.ld * sizeof(double),
                                        w * sizeof(double), rows, hipMemcpyDeviceToHost, c->stream));
    } else {
        QPAD_HIP_CHECK(hipMemcpy2DAsync(dev, c->g.ld * sizeof(double), host_in, w * sizeof(double),
                                        w * sizeof(double), rows, hipMemcpyHostToDevice, c->stream));
    }
    return sync(c);
}

int qpad_get_grid(qpad_ctx *c, qpad_grid *out)
{
    if (!c || !out) return fail(QPAD_E_INVALID, "null argument");
    const Grid &g = c->g;
    out->nr = g.nr;
    out->nxi = g.nxi;
    out->max_mode = g.M;
    out->dr = g.dr;
    out->dxi = g.dxi;
    out->relax_fac = g.relax_fac;
    out->bnd = g.bnd;
    out->iter_max = g.iter_max;
    out->iter_reltol = g.reltol;
    out->iter_abstol = g.abstol;
    out->dt = g.dt;
    return 0;
}

int qpad_get_field(qpad_ctx *c, int32_t fid, double *out)
{
    if (!c || !valid_fid(fid) || !out) return fail(QPAD_E_INVALID, "bad field id");
    return copy_rows(c, c->f[fid], c->ncomp[fid], out, nullptr);
}

int qpad_set_field(qpad_ctx *c, int32_t fid, const double *in)
{
    if (!c || !valid_fid(fid) || !in) return fail(QPAD_E_INVALID, "bad field id");
    return copy_rows(c, c->f[fid], c->ncomp[fid], nullptr, in);
}

int qpad_get_hist(qpad_ctx *c, int32_t fid, int32_t s, double *out)
{
    int nc = 0;
    double *h = c ? hist_ptr(c, fid, &nc) : nullptr;
    if (!h || !out || s < 0 || s > c->g.nxi + 1) return fail(QPAD_E_INVALID, "bad history request");
    return copy_rows(c, h + (size_t)s * f1_size(c, nc), nc, out, nullptr);
}

int qpad_set_hist(qpad_ctx *c, int32_t fid, int32_t s, const double *in)
{
    int nc = 0;
    double *h = c ? hist_ptr(c, fid, &nc) : nullptr;
    if (!h || !in || s < 0 || s > c->g.nxi + 1) return fail(QPAD_E_INVALID, "bad history request");
    return copy_rows(c, h + (size_t)s * f1_size(c, nc), nc, nullptr, in);
}

int qpad_species_count(qpad_ctx *c, int32_t s, int64_t *npp)
{
    if (!c || s < 0 || s >= c->nspecies || !npp) return fail(QPAD_E_INVALID, "bad species");
    TRY(refresh_npp(c, s));
    *npp = c->sp[s].npp;
    return 0;
}

int qpad_get_species(qpad_ctx *c, int32_t s, double *x1, double *x2, double *p1, double *p2, double *p3,
                     double *q, double *gamma, double *psi, int32_t *id)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    TRY(refresh_npp(c, s));
    Species &sp = c->sp[s];
    const size_t n = (size_t)sp.npp;
    double *outs[8] = {x1, x2, p1, p2, p3, q, gamma, psi};
    const double *srcs[8] = {sp.x1, sp.x2, sp.p1, sp.p2, sp.p3, sp.q, sp.gamma, sp.psi};
    for (int f = 0; f < 8; f++)
        if (outs[f] && n) QPAD_HIP_CHECK(hipMemcpyAsync(outs[f], srcs[f], n * sizeof(double),
                                                        hipMemcpyDeviceToHost, c->stream));
    if (id && n) QPAD_HIP_CHECK(hipMemcpyAsync(id, sp.id, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int qpad_set_species(qpad_ctx *c, int32_t s, int64_t n, const double *x1, const double *x2, const double *p1,
                     const double *p2, const double *p3, const double *q, const double *gamma, const double *psi,
                     const int32_t *id)
{
    if (!c || s < 0 || s >= c->nspecies) return fail(QPAD_E_INVALID, "bad species");
    Species &sp = c->sp[s];
    if (n < 0 || n > sp.npmax) return fail(QPAD_E_INVALID, "too many particles for this species");
    if (!x1 || !x2 || !p1 || !p2 || !p3 || !q) return fail(QPAD_E_INVALID, "x, p and q are required");
    const double *srcs[8] = {x1, x2, p1, p2, p3, q, gamma, psi};
    double *dsts[8] = {sp.x1, sp.x2, sp.p1, sp.p2, sp.p3, sp.q, sp.gamma, sp.psi};
    for (int f = 0; f < 8; f++) {
        if (!n) break;
        if (srcs[f]) QPAD_HIP_CHECK(hipMemcpyAsync(dsts[f], srcs[f], n * sizeof(double),
                                                   hipMemcpyHostToDevice, c->stream));
        else TRY(zero_dev(c, dsts[f], (size_t)n));
    }
    std::vector<int32_t> ids;
    const int32_t *idp = id;
    if (!idp) {
        ids.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) ids[(size_t)i] = (int32_t)i;
        idp = ids.data();
    }
    if (n) QPAD_HIP_CHECK(hipMemcpyAsync(sp.id, idp, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    TRY(sync(c));
    sp.npp = n;
    TRY(launch_keys(c, s));
    TRY(launch_sort(c, s));
    TRY(refresh_npp(c, s));
    return 0;
}

int qpad_get_species_qn(qpad_ctx *c, int32_t s, double *out)
{
    if (!c || s < 0 || s >= c->nspecies || !out) return fail(QPAD_E_INVALID, "bad species");
    return copy_rows(c, c->sp[s].qn, 1, out, nullptr);
}

int qpad_set_species_qn(qpad_ctx *c, int32_t s, const double *in)
{
    if (!c || s < 0 || s >= c->nspecies || !in) return fail(QPAD_E_INVALID, "bad species");
    return copy_rows(c, c->sp[s].qn, 1, nullptr, in);
}

int qpad_beam_count(qpad_ctx *c, int32_t b, int64_t *npp)
{
    if (!c || b < 0 || b >= c->nbeams || !npp) return fail(QPAD_E_INVALID, "bad beam");
    TRY(launch_beam_sort(c, b));
    TRY(refresh_beam_npp(c, b));
    *npp = c->bm[b].npp;
    return 0;
}

int qpad_get_beam(qpad_ctx *c, int32_t b, double *x1, double *x2, double *x3, double *p1, double *p2, double *p3,
                  double *q)
{
    if (!c || b < 0 || b >= c->nbeams) return fail(QPAD_E_INVALID, "bad beam");
    TRY(launch_beam_sort(c, b));
    TRY(refresh_beam_npp(c, b));
    Beam &bm = c->bm[b];
    const size_t n = (size_t)bm.npp;
    double *outs[7] = {x1, x2, x3, p1, p2, p3, q};
    const double *srcs[7] = {bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q};
    for (int f = 0; f < 7; f++)
        if (outs[f] && n) QPAD_HIP_CHECK(hipMemcpyAsync(outs[f], srcs[f], n * sizeof(double),
                                                        hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int qpad_set_beam(qpad_ctx *c, int32_t b, int64_t n, const double *x1, const double *x2, const double *x3,
                  const double *p1, const double *p2, const double *p3, const double *q)
{
    if (!c || b < 0 || b >= c->nbeams) return fail(QPAD_E_INVALID, "bad beam");
    Beam &bm = c->bm[b];
    if (n < 0 || n > bm.npmax) return fail(QPAD_E_INVALID, "too many beam particles");
    if (!x1 || !x2 || !x3 || !p1 || !p2 || !p3 || !q) return fail(QPAD_E_INVALID, "null beam array");
    const double *srcs[7] = {x1, x2, x3, p1, p2, p3, q};
    double *dsts[7] = {bm.x1, bm.x2, bm.x3, bm.p1, bm.p2, bm.p3, bm.q};
    for (int f = 0; f < 7 && n; f++)
        QPAD_HIP_CHECK(hipMemcpyAsync(dsts[f], srcs[f], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TRY(sync(c));
    bm.npp = n;
    TRY(launch_beam_sort(c, b));
    TRY(refresh_beam_npp(c, b));
    return 0;
}

int qpad_kernel_time(qpad_ctx *c, int32_t which, double *ms, int64_t *launches)
{
    if (!c || which < 0 || which > 2) return fail(QPAD_E_INVALID, "bad timer id");
    pool_drain(c, c->pools[which]);
    if (ms) *ms = c->pools[which].acc_ms;
    if (launches) *launches = c->pools[which].launches;
    return 0;
}

int qpad_kernel_time_reset(qpad_ctx *c, int32_t enable)
{
    if (!c) return fail(QPAD_E_INVALID, "null context");
    TRY(sync(c));
    for (auto &p : c->pools) {
        pool_drain(c, p);
        p.acc_ms = 0.0;
        p.launches = 0;
    }
    c->timing = enable != 0;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched instances: several independent contexts stepped in one call
// ---------------------------------------------------------------------------
struct qpad_batch {
    std::vector<qpad_ctx *> ctx;
    int graph = 0;   // mode: 0 eager (host thread per context), 1 graph, 2 eager (one host thread)
    // per context, its captured slice loops.  The kernels capture the
    // particle buffers by address, and the buffers swap once per sort: an
    // even number of times in a slice loop, once more in the step's tail
    // (the renewed plasma is sorted).  So the step alternates between two
    // buffer assignments, and each context keeps one graph per assignment,
    // keyed by the first species' x1 buffer at capture.
    std::vector<hipGraphExec_t> exec[2];
    std::vector<const void *> key[2];
};

namespace qpad {

constexpr int kBatchMaxThreads = 64;   // host threads of an eager batch step

// the slice loop of one 3D step of c, captured into a graph: all iter_max
// iterations queued (no host feedback), no timing events, no seed network.
// The particle buffers swap once per slice, so an even nxi returns them to
// the captured state after every replay.
static int capture_slices(qpad_ctx *c, hipGraphExec_t *exec)
{
    if (c->g.nxi % 2) return fail(QPAD_E_INVALID, "graph mode needs an even nxi");
    const int adaptive = c->adaptive;
    const bool timing = c->timing;
    c->adaptive = 0;
    c->timing = false;
    hipGraph_t g = nullptr;
    QPAD_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    int rc = 0;
    for (int j = 1; j <= c->g.nxi && !rc; j++) rc = slice_enqueue(c, j, false);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->adaptive = adaptive;
    c->timing = timing;
    if (rc) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc;
    }
    QPAD_HIP_CHECK(e);
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    QPAD_HIP_CHECK(ei);
    return 0;
}

static void batch_drop_graphs(qpad_batch *b)
{
    for (int k = 0; k < 2; k++) {
        for (auto &x : b->exec[k])
            if (x) (void)hipGraphExecDestroy(x);
        b->exec[k].assign(b->ctx.size(), nullptr);
        b->key[k].assign(b->ctx.size(), nullptr);
    }
}

}  // namespace qpad

extern "C" {

int qpad_batch_create(qpad_ctx *const *ctxs, int32_t n, qpad_batch **out)
{
    if (!ctxs || !out || n < 1) return fail(QPAD_E_INVALID, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return fail(QPAD_E_INVALID, "null context in batch");
        for (int k = 0; k < i; k++)
            if (ctxs[k] == ctxs[i]) return fail(QPAD_E_INVALID, "a context appears twice in the batch");
        if (ctxs[i]->device != ctxs[0]->device) return fail(QPAD_E_INVALID, "batch contexts on different devices");
    }
    qpad_batch *b = new qpad_batch();
    b->ctx.assign(ctxs, ctxs + n);
    batch_drop_graphs(b);
    *out = b;
    return 0;
}

int qpad_batch_destroy(qpad_batch *b)
{
    if (!b) return 0;
    for (qpad_ctx *c : b->ctx) (void)hipStreamSynchronize(c->stream);
    batch_drop_graphs(b);
    delete b;
    return 0;
}

// tune every context's launch shape (qpad_tune_launch); captured graphs are
// dropped (they hold the old shapes)
int qpad_batch_tune(qpad_batch *b, int32_t reps)
{
    if (!b) return fail(QPAD_E_INVALID, "null batch");
    for (qpad_ctx *c : b->ctx) TRY(qpad_tune_launch(c, reps, nullptr));
    batch_drop_graphs(b);
    return 0;
}

int qpad_batch_mode(qpad_batch *b, int32_t mode)
{
    if (!b) return fail(QPAD_E_INVALID, "null batch");
    if (mode < 0 || mode > 2) return fail(QPAD_E_INVALID, "batch mode must be 0, 1 or 2");
    if ((mode == 1) != (b->graph == 1)) batch_drop_graphs(b);
    b->graph = mode;
    return 0;
}

// One 3D step of every context: the heads, then the slices of all contexts
// (each context on its own stream, so their kernels run side by side: queued
// by an event-driven loop over the contexts, or as each context's captured
// slice graph), then the tails.  Every context's results equal its own
// qpad_step3d.
int qpad_batch_step3d(qpad_batch *b, int32_t *total_iters_out)
{
    if (!b) return fail(QPAD_E_INVALID, "null batch");
    const int n = (int)b->ctx.size();
    // the seed network's stream joins a context's slice loop through events
    // the eager modes queue like qpad_step3d does; a captured slice graph
    // holds no such second stream
    for (qpad_ctx *c : b->ctx)
        if (c->subcyc) return fail(QPAD_E_STATE, "batches step without sub-cycling (its slices wait for the host)");
    if (b->graph == 1)
        for (qpad_ctx *c : b->ctx) {
            if (seednet_active(c))
                return fail(QPAD_E_STATE, "graph batches step without the seed network (use an eager mode)");
            for (int s = 0; s < c->nspecies; s++)
                if (c->sp[s].neutral)
                    return fail(QPAD_E_STATE, "graph batches step without neutrals: each slice's release "
                                              "count comes back to the host (use an eager mode)");
        }
    for (qpad_ctx *c : b->ctx) {
        TRY(step_begin_enqueue(c));
        memset(c->it_host, 0, sizeof(IterState));
        QPAD_HIP_CHECK(hipMemcpyAsync(c->it, c->it_host, sizeof(IterState), hipMemcpyHostToDevice, c->stream));
    }
    if (b->graph == 1) {
        for (int i = 0; i < n; i++) {
            qpad_ctx *c = b->ctx[i];
            const void *key = c->nspecies > 0 ? (const void *)c->sp[0].x1 : nullptr;
            int k = -1;
            for (int q = 0; q < 2 && k < 0; q++)
                if (b->exec[q][i] && b->key[q][i] == key) k = q;
            if (k < 0) {
                k = b->exec[0][i] ? 1 : 0;
                if (b->exec[k][i]) (void)hipGraphExecDestroy(b->exec[k][i]);
                b->exec[k][i] = nullptr;
                TRY(capture_slices(c, &b->exec[k][i]));
                b->key[k][i] = key;
            }
            QPAD_HIP_CHECK(hipGraphLaunch(b->exec[k][i], c->stream));
        }
    } else if (b->graph == 0) {
        // one host thread per context (each runs the context's slice loop
        // with its own iteration policy, waiting on its own decisions), so
        // that the launches of all contexts are issued in parallel.  The
        // kernels each context queues are the same as in its own qpad_step3d.
        std::vector<int> rcs(n, 0);
        std::vector<std::string> msgs(n);
        std::atomic<int> next_ctx{0};
        auto worker = [&]() {
            for (;;) {
                const int i = next_ctx.fetch_add(1);
                if (i >= n) break;
                qpad_ctx *c = b->ctx[i];
                int rc = hipSetDevice(c->device) == hipSuccess ? 0 : fail(QPAD_E_HIP, "hipSetDevice failed");
                for (int j = 1; j <= c->g.nxi && !rc; j++) rc = slice_enqueue(c, j, true);
                if (rc) {
                    rcs[i] = rc;
                    msgs[i] = qpad_last_error();
                }
            }
        };
        const int nth = std::min(n, kBatchMaxThreads);
        std::vector<std::thread> pool;
        for (int k = 0; k < nth; k++) pool.emplace_back(worker);
        for (auto &th : pool) th.join();
        for (int i = 0; i < n; i++)
            if (rcs[i]) return fail(rcs[i], msgs[i]);
    } else {
        // Single host thread: each context runs its own slice loop with its
        // own iteration policy.  The host never waits on one context: it
        // polls every context's pending decision (dec_ev) and serves the
        // ready ones, so that all streams stay fed.
        std::vector<SliceRun> st(n);
        std::vector<int> next(n, 1);
        int active = 0;
        for (int i = 0; i < n; i++) {
            qpad_ctx *c = b->ctx[i];
            if (c->g.nxi < 1) continue;
            active++;
            TRY(slice_start(c, next[i]++, true, st[i]));
        }
        while (active > 0) {
            for (int i = 0; i < n; i++) {
                qpad_ctx *c = b->ctx[i];
                if (st[i].j == 0 && next[i] > c->g.nxi) continue;
                if (st[i].waiting) {
                    const hipError_t q = hipEventQuery(c->dec_ev);
                    if (q == hipErrorNotReady) continue;
                    QPAD_HIP_CHECK(q);
                    TRY(slice_decided(c, st[i]));
                }
                // a slice that finished starts the next one at once
                while (!st[i].waiting && st[i].j == 0 && next[i] <= c->g.nxi)
                    TRY(slice_start(c, next[i]++, true, st[i]));
                if (st[i].j == 0 && next[i] > c->g.nxi) active--;
            }
        }
    }
    for (int i = 0; i < n; i++) {
        qpad_ctx *c = b->ctx[i];
        TRY(read_iter_state(c));
        if (total_iters_out) total_iters_out[i] = c->it_host->total_iters;
        TRY(qpad_step3d_end(c));
    }
    return 0;
}

}  // extern "C"

