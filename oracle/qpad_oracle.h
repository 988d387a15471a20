/*
 * qpad_oracle.h -- serial FP64 CPU restatement of QPAD's per-xi-slice plasma
 * response and once-per-3D-step beam push.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the correctness oracle (and the
 * timed CPU baseline) for the HIP library in qpad_amd/csrc.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may link or call
 * it, and only as the checker.  The product never routes through it.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the QPAD source tree).  Data layout (shared with the HIP library so the two
 * can be compared array-for-array):
 *   slice field  f[(c*nm + k)*ld + i]   c = component, k = mode array,
 *                                        i = radial node 0..nr+1 (0 and nr+1
 *                                        are the guard nodes of ufield%f1)
 *   mode arrays  k = 0 : Re m=0 ;  k = 2m-1 : Re m ;  k = 2m : Im m
 *   history      h[(s*ncomp + c)*nm + k)*ld + i], slice s = 0..nxi+1
 *                (s = 1..nxi are QPAD's f2 slices, s = nxi+1 is the upper
 *                 guard slice, s = 0 is unused)
 *
 * Parity pins (tests/test_oracle_pins.py, tests/oracle_pins.py): every
 * routine is checked against what the mathematics fixes -- closed forms,
 * polynomial exactness of each stencil (axis and edge rows included),
 * manufactured solutions with second-order convergence, vacuum fall-off of
 * the boundary folds, invariants, dense solves, known-answer vectors -- and
 * tests/oracle_mutants.py shows each check failing on a copy of this file
 * with one plausible mistake planted.  Restated without a mathematical pin:
 * the B- outer fold using m + 1 (field_solver_class.f03:532-540), the m = 0
 * zero ghost of psi/Ez/Bz, and the order of the slice driver (checked only
 * by the linear-wake test).  HYPRE StructCycRed (the reference's
 * tridiagonal solver, third-party, unversioned) is replaced by Thomas'
 * algorithm on the same matrices: parity unpinned against HYPRE's rounding.
 * gfortran's random_number stream is replaced by Philox4x32-10: thermal
 * plasma (uth != 0) and random beams are parity unpinned against QPAD's own
 * draws; beams are shared with the GPU side as particle arrays instead.
 */
#ifndef QPAD_ORACLE_H
#define QPAD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field kinds of field_solver (source/param.f03:75-84) */
enum { ORA_FK_PSI = 0, ORA_FK_EZ = 1, ORA_FK_BZ = 2, ORA_FK_BT = 3,
       ORA_FK_BPLUS = 4, ORA_FK_BMINUS = 5 };
/* boundary conditions (source/param.f03:71-74) */
enum { ORA_BND_ZERO = 2, ORA_BND_OPEN = 3 };

/* field ids of the slice state (source/sim_fields_class.f03:147-160) */
enum { ORA_PSI = 0, ORA_E, ORA_E_SPE, ORA_B, ORA_B_SPE, ORA_B_BEAM,
       ORA_CU, ORA_ACU, ORA_AMU, ORA_DCU, ORA_Q_SPE, ORA_Q_BEAM, ORA_NFIELD };

typedef struct {
    int nr, nxi, max_mode;     /* grid(1), grid(2), simulation.max_mode     */
    double dr, dxi;            /* options_class.f03:89,93                   */
    double relax_fac;          /* sim_fields_class.f03:137-140              */
    int bnd;                   /* ORA_BND_OPEN / ORA_BND_ZERO               */
    int iter_max;              /* simulation.iter_max                       */
    double iter_reltol, iter_abstol;
    double dt;                 /* 3D time step (beam push)                  */
} ora_grid;

typedef struct {
    int ppc1, ppc2, num_theta; /* species(i).ppc, num_theta                 */
    double qm, mass;           /* species(i).q, species(i).m                */
    double density;
    double uth[3];
    int neutralized;           /* species(i).neutralized (default true)     */
    double den_min;            /* species(i).den_min (default 1e-10)        */
    uint64_t seed;             /* Philox key for thermal draws (uth != 0)   */
    int push_type;             /* species(i).push_type: ORA_PUSH2_STD
                                  ("standard"), ORA_PUSH2_ROBUST, ORA_PUSH2_
                                  STD_PGC, ORA_PUSH2_ROBUST_PGC (param.f03:
                                  63-64, sim_plasma_class.f03:123-137)       */
} ora_species_cfg;

#define ORA_PUSH2_STD 0
#define ORA_PUSH2_ROBUST 1
#define ORA_PUSH2_STD_PGC 4
#define ORA_PUSH2_ROBUST_PGC 5

typedef struct {
    int64_t total_num;         /* beam(i).total_num                         */
    double total_charge;       /* beam(i).total_charge                      */
    double qm, mass, gamma;
    double center[3], sigma[3];/* gauss_center, gauss_sigma (xi relative to
                                  the box start, i.e. already minus z0)     */
    double range_lo[3], range_hi[3];
    double uth[3];
    int quiet;                 /* quiet_start                               */
    int evol;                  /* evolution                                 */
    uint64_t seed;
} ora_beam_cfg;

/* beam(i) section of profile_type "standard" (fdist3d_std_class.f03:
   150-436): profiles per direction 0 uniform, 1 gaussian, 2 piecewise-
   linear; geometry 0 cartesian, 1 cylindrical; xi positions relative to the
   box start.  Same layout as the library's qpad_beam_std_cfg. */
typedef struct {
    int ppc[3];
    int num_theta;
    int geometry;
    int profile[3];
    double qm, mass, density, gamma, den_min;
    double gauss_center[3], gauss_sigma[3];
    int pw_n[3];
    double pw_x[3][32], pw_fx[3][32];
    double range_lo[3], range_hi[3];
    double uth[3];
    double alpha[2], twiss_ctr[2], twiss_beta[2];
    int offset_n[2];
    double offset_xi0[2];
    double offset_c[2][8];
    int quiet;
    int evol;
    uint64_t seed;
} ora_beam_std_cfg;

typedef struct {
    int64_t npp, npmax;
    double *x1, *x2, *p1, *p2, *p3, *q, *gamma, *psi;
    int32_t *id;
    double qbm, qm, dt;
    ora_species_cfg cfg;
    double *q_own, *qn;       /* species%q, species%qn slice arrays        */
    double *cu_own, *acu_own, *amu_own;
    double *q_hist;           /* species q history (cbq)                   */
    /* a neutral gas (neutral_class.f03): the particles are the electrons
       ionisation has released; qn is the ion charge rho_ion it has left */
    int is_neutral;
    int multi_max;            /* ionisation levels followed               */
    double wp;                /* plasma frequency (SI) of simulation.n0    */
    const double *rate;       /* [multi_max][3] ADK (C1, C2, C3) per level */
    double *multi_ion;        /* [n_theta * nr][multi_max + 2], cell
                                 c = k * nr + j (theta bin k, radial cell
                                 j): levels 1..multi_max, the neutral
                                 residue, the discrete total ion charge    */
    double *ion_old;          /* [n_theta * nr]                            */
    int *ppc_add;             /* [n_theta * nr] electrons released          */
    double *add_x1, *add_x2, *add_q;   /* part_add: the released electrons'
                                          positions and ion charges        */
    int64_t n_add, n_added;   /* in part_add; released since the renew     */
    double *rho_hist;         /* rho_ion history [(nxi + 2)][nm][ld]       */
    /* a non-uniform plasma: the transverse density at every lattice point
       [num_theta][nr][ppc1][ppc2] (NULL: uniform) and the longitudinal
       factor of the next injection */
    double *den_perp;
    double den_lon;
} ora_species;

/* neutrals(i) section (sim_plasma_class.f03:148-176, fdist2d for the
   'neutrals' section): uniform profile, element by atomic number (H, He,
   Li, C, N, O, Ne, Ar), ion_max levels at most, pusher; n0 of the
   simulation section gives the plasma frequency */
typedef struct {
    int ppc1, ppc2, num_theta;
    double qm, mass, density, den_min;
    int element, ion_max, push_type;
    double n0;
} ora_neutral_cfg;

typedef struct {
    int64_t npp, npmax;
    double *x1, *x2, *x3, *p1, *p2, *p3, *q;
    int32_t *id;
    double qbm, dt;
    int evol;
    double *q_own_hist;       /* beam3d%q f2                               */
} ora_beam;

typedef struct {
    double *lo, *inv, *cp;    /* Thomas factors, n = nr rows               */
} ora_factor;

typedef struct {
    ora_grid g;
    int nm, ld;
    double *f[ORA_NFIELD];    /* slice arrays                              */
    int ncomp[ORA_NFIELD];
    double *h_psi, *h_e, *h_b, *h_b_spe, *h_e_spe, *h_cu, *h_q_spe, *h_q_beam;
    int nspecies, nbeams;
    ora_species sp[8];
    ora_beam bm[8];
    /* factors[kind][mode] */
    ora_factor fac[6][16];
    /* convergence_tester saved vectors (simulation_class.f03:534) */
    double *conv_re, *conv_im;
    /* diagnostics of the last slice */
    int last_iters;
    double last_rel, last_abs;
    /* laser (PGC): the envelope history (cfr, cfi of field_laser, slices
       -1 .. nxi+1, [(nxi + 3)][nm][ld]), the slice's envelope and gradients
       (laser_all: ar, ai [nm][ld]; ar_grad, ai_grad [3][nm][ld], components
       r, phi, z) and the plasma susceptibility chi with its history
       [(nxi + 2)][nm][ld] */
    int laser_on;
    double *lz_ar, *lz_ai;
    double *ar, *ai, *ar_grad, *ai_grad;
    double *chi, *h_chi;
    /* sub-cycling (proj_subcyc): on, expansion_fac_max,
       expansion_fac_clamped, dt_2d_min; the last slice's sub-cycles, their
       2D step and the expansion factor they came from */
    int subcyc;
    double sc_fac_max, sc_fac_clamped, sc_dt_min;
    int last_nsub;
    double last_dt_sub, last_exp_fac;
} ora_sim;

/* ---- low-level building blocks (exported for unit tests) ---- */
void ora_matrix(int kind, int mode, int nr, double dr, double relax_fac,
                int bnd, double *a, double *b, double *c);
void ora_thomas_factor(int n, const double *a, const double *b,
                       const double *c, double *lo, double *inv, double *cp);
void ora_thomas_solve(int n, const double *lo, const double *inv,
                      const double *cp, double *d);

void ora_solve_psi(const ora_sim *s, const double *q, double *psi);
void ora_solve_ez(const ora_sim *s, const double *jay, double *e);
void ora_solve_et(const ora_sim *s, const double *b, const double *psi,
                  double *e);
void ora_solve_bz(const ora_sim *s, const double *jay, double *b);
void ora_solve_bt_beam(const ora_sim *s, const double *q, double *b);
void ora_solve_bt_iter(const ora_sim *s, const double *djdxi,
                       const double *jay, double *b);
void ora_solve_djdxi(const ora_sim *s, const double *acu, const double *amu,
                     double *dcu);
void ora_conv_record(ora_sim *s, const double *b);
void ora_conv_compare(ora_sim *s, const double *b, double *rel, double *abs_);

void ora_qdeposit(const ora_sim *s, const ora_species *sp, double *q);
void ora_amjdeposit_robust(const ora_sim *s, ora_species *sp,
                           const double *e, const double *b, double *cu,
                           double *amu, double *dcu, double dt);
void ora_push_u_robust(const ora_sim *s, ora_species *sp, const double *e,
                       const double *b, double dt);
void ora_interp_psi(const ora_sim *s, ora_species *sp, const double *psi);
void ora_amjdeposit_std(const ora_sim *s, ora_species *sp,
                        const double *e, const double *b, double *cu,
                        double *amu, double *dcu, double dt);
void ora_push_u_std(const ora_sim *s, ora_species *sp, const double *e,
                    const double *b, double dt);
void ora_push_x(ora_species *sp, double dt);
/* laser (PGC) */
int ora_set_laser(ora_sim *s, const double *ar_hist, const double *ai_hist);
void ora_laser_slice(ora_sim *s, int j);
void ora_amjdeposit_pgc(const ora_sim *s, ora_species *sp, const double *e, const double *b,
                        double *cu, double *amu, double *dcu, double dt, int std_);
void ora_push_u_pgc(const ora_sim *s, ora_species *sp, const double *e, const double *b, double dt);
void ora_deposit_chi(const ora_sim *s, const ora_species *sp, double *chi);
void ora_update_bound(const ora_sim *s, ora_species *sp);
/* neutrals */
int ora_add_neutral(ora_sim *s, const ora_neutral_cfg *cfg);
/* ADK coefficients (C1, C2, C3) of level `level` (1-based) of element Z;
   NULL where the table has none */
const double *ora_adk(int element, int level);
void ora_neutral_ionize(const ora_sim *s, ora_species *sp, const double *e);
void ora_neutral_add(const ora_sim *s, ora_species *sp);
void ora_neutral_ion_deposit(const ora_sim *s, ora_species *sp);
void ora_neutral_renew(const ora_sim *s, ora_species *sp);
/* sub-cycling */
void ora_set_subcycling(ora_sim *s, int on, double exp_fac_max, double exp_fac_clamped, double dt_2d_min);
double ora_exp_fac_max(const ora_sim *s);
int ora_subcyc_step(double exp_fac, double exp_fac_max, double dt, double dt_min, double *dt_sub);
void ora_clamp_exp_fac(ora_species *sp, double exp_fac_clamped);
void ora_last_subcycles(const ora_sim *s, int *n, double *dt_sub, double *exp_fac);
void ora_interp_point(const ora_sim *s, const double *f, int ncomp,
                      double x1, double x2, double *out);

/* ---- simulation object ---- */
ora_sim *ora_create(const ora_grid *g);
void ora_destroy(ora_sim *s);
int ora_add_species(ora_sim *s, const ora_species_cfg *cfg);
int ora_add_beam(ora_sim *s, const ora_beam_cfg *cfg, int64_t npmax);
int ora_add_beam_std(ora_sim *s, const ora_beam_std_cfg *cfg);
/* the lattice beam with the density of every point of the box xi cells
   k0 .. k0 + nk - 1, r cells 1 .. ni given (without cfg's density factor) */
int ora_add_beam_std_table(ora_sim *s, const ora_beam_std_cfg *cfg, const double *den, int k0, int nk, int ni);
int ora_beam_set(ora_sim *s, int ib, int64_t n, const double *x1,
                 const double *x2, const double *x3, const double *p1,
                 const double *p2, const double *p3, const double *q);
void ora_species_inject(ora_sim *s, int isp);
/* the profile of the species' next injections (den_perp NULL: uniform) */
int ora_species_set_profile(ora_sim *s, int isp, const double *den_perp, double den_lon);
void ora_species_init_q(ora_sim *s, int isp);
void ora_beam_inject(ora_sim *s, int ib, const ora_beam_cfg *cfg);
void ora_beam_qdeposit(ora_sim *s, int ib, double *q_hist);
void ora_beam_push(ora_sim *s, int ib);

void ora_step3d_begin(ora_sim *s);
void ora_slice(ora_sim *s, int j, int *n_iter);
void ora_step3d_end(ora_sim *s, int istep);
void ora_run(ora_sim *s, int nstep3d, int nslices);

double *ora_field(ora_sim *s, int fid);
double *ora_hist(ora_sim *s, int fid);
/* laser arrays (NULL before ora_set_laser or a PGC species): 0 ar, 1 ai,
   2 ar_grad, 3 ai_grad, 4 chi, 5 h_chi, 6 the ar history, 7 the ai history */
double *ora_laser_array(ora_sim *s, int which);
int ora_nm(const ora_sim *s);
int ora_ld(const ora_sim *s);
ora_species *ora_get_species(ora_sim *s, int isp);
ora_beam *ora_get_beam(ora_sim *s, int ib);
int ora_last_iters(const ora_sim *s);
/* the relative and absolute residuals of the last slice's final 'compare' */
void ora_last_residuals(const ora_sim *s, double *rel, double *abs_);

/* timers by QPAD's event names (sysutil_module.f03:307-330; 'set source'
   is field_src_class.f03:293): seconds accumulated since ora_timing(1) */
enum { ORA_T_SOLVE_PSI = 0, ORA_T_SOLVE_BZ, ORA_T_SOLVE_BEAM_BT, ORA_T_SOLVE_PLASMA_BT,
       ORA_T_SOLVE_EZ, ORA_T_SOLVE_PLASMA_ET, ORA_T_SET_SOURCE, ORA_T_COPY_SLICES,
       ORA_T_ARITHMETICS, ORA_T_DEPOSIT_2D, ORA_T_PUSH_2D, ORA_T_MOVE_2D,
       ORA_T_DEPOSIT_3D, ORA_T_PUSH_3D, ORA_NTIMER };
void ora_timing(int enable);
void ora_timers(double *out);

/* Philox4x32-10 (Salmon et al., SC'11) and the Gaussian draw built on it */
void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]);
void ora_normal4(uint64_t seed, uint64_t stream, uint32_t sub,
                 double out[4]);

#ifdef __cplusplus
}
#endif
#endif
