/*
 * qpad_hip.h -- C ABI of libqpad_hip.so, the MI355X (gfx950) implementation of
 * QPAD's per-xi-slice plasma response and per-3D-step beam push.
 *
 * The entry points are the type-bound procedures QPAD's simulation loop
 * calls (source/simulation_class.f03:226-520), flattened to plain C so that a
 * Fortran ISO_C_BINDING module (qpad_amd/fortran/qpad_hip_mod.f03) can bind
 * them one-for-one.  Each declaration cites the reference procedure it
 * replaces (paths relative to the QPAD source tree).
 *
 * Conventions
 *  - every function returns int: 0 = success, < 0 = error (see QPAD_E_*);
 *    qpad_last_error() returns a message for the calling thread.  The
 *    library never aborts and never prints.
 *  - all particle and field state lives in device memory inside the context;
 *    host arrays cross only in the qpad_get_* / qpad_set_* calls.
 *  - slice field arrays cross as [ncomp][2*max_mode+1][nr+2] doubles: mode
 *    array k = 0 is Re m=0, k = 2m-1 is Re m, k = 2m is Im m; node i = 0..nr+1
 *    is QPAD's ufield%f1(:, i) (0 and nr+1 are guard nodes).
 *  - slice indices j are 1-based like QPAD's (1..nxi); history slice nxi+1 is
 *    the upper guard slice of ufield%f2.
 *  - one context = one simulation instance on the current HIP device; calls on
 *    one context are serialised on its stream and return when the work is
 *    done (synchronous at the boundary, like the Fortran procedures).
 */
#ifndef QPAD_HIP_H
#define QPAD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPAD_HIP_ABI_VERSION 2

/* error codes */
#define QPAD_OK 0
#define QPAD_E_INVALID (-1)   /* bad argument / unsupported configuration   */
#define QPAD_E_HIP (-2)       /* a HIP runtime call failed                  */
#define QPAD_E_NOMEM (-3)     /* device allocation failed                   */
#define QPAD_E_STATE (-4)     /* call out of sequence                       */

/* field ids (source/sim_fields_class.f03:147-160) */
#define QPAD_F_PSI 0
#define QPAD_F_E 1
#define QPAD_F_E_SPE 2
#define QPAD_F_B 3
#define QPAD_F_B_SPE 4
#define QPAD_F_B_BEAM 5
#define QPAD_F_CU 6
#define QPAD_F_ACU 7
#define QPAD_F_AMU 8
#define QPAD_F_DCU 9
#define QPAD_F_Q_SPE 10
#define QPAD_F_Q_BEAM 11
/* the laser envelope of the slice and its gradients (field_laser laser_all:
   cfr, cfi and ar_grad, ai_grad with components r, phi, z;
   field_laser_class.f03:952-1012, :637-750) and the plasma susceptibility
   chi (sim_lasers%chi, sim_lasers_class.f03:176-200); chi has a history */
#define QPAD_F_AR 12
#define QPAD_F_AI 13
#define QPAD_F_AR_GRAD 14
#define QPAD_F_AI_GRAD 15
#define QPAD_F_CHI 16
#define QPAD_NFIELD 17

/* boundary conditions (source/param.f03:71-74) */
#define QPAD_BND_ZERO 2
#define QPAD_BND_OPEN 3

/* plasma pusher (source/param.f03:63-64, sim_plasma_class.f03:123-137):
   "standard", "robust", "standard_pgc", "robust_pgc" (the ponderomotive
   guiding centre pushers, which push in the laser envelope of
   qpad_set_laser; a zero envelope until one is set) */
#define QPAD_PUSH2_STD 0
#define QPAD_PUSH2_ROBUST 1
#define QPAD_PUSH2_STD_PGC 4
#define QPAD_PUSH2_ROBUST_PGC 5

/* grid and iteration settings: options_class.f03:44-117 (nr, nxi, dr, dxi),
   simulation_class.f03:141-172 (iter_max/reltol/abstol, dt),
   sim_fields_class.f03:112-140 (field_boundary, relax_fac) */
typedef struct {
    int32_t nr, nxi, max_mode;
    double dr, dxi;
    double relax_fac;
    int32_t bnd;
    int32_t iter_max;
    double iter_reltol, iter_abstol;
    double dt;
} qpad_grid;

/* species(i) section: fdist2d_class.f03:117-272 (uniform profiles only) and
   sim_plasma_class.f03:116-157 (q, m, push_type) */
typedef struct {
    int32_t ppc1, ppc2, num_theta;
    double qm, mass, density;
    double uth[3];
    int32_t neutralized;
    double den_min;
    uint64_t seed;
    int32_t push_type;
} qpad_species_cfg;

/* beam(i) section for profile_type "random", Cartesian Gaussian profile:
   fdist3d_rnd_class.f03:104-356; xi positions are relative to the box start */
typedef struct {
    int64_t total_num;
    double total_charge;
    double qm, mass, gamma;
    double center[3], sigma[3];
    double range_lo[3], range_hi[3];
    double uth[3];
    int32_t quiet;
    int32_t evol;
    uint64_t seed;
} qpad_beam_cfg;

/* beam(i) section for profile_type "standard" (the default): the lattice
   beam of fdist3d_std_class.f03:150-436 (settings) and :439-613
   (injection).  Per direction a density profile (fdist3d_std_lib.f03:
   uniform, gaussian exp(-(x - c)^2 / (2 s^2)), piecewise-linear in pw_x /
   pw_fx), applied to x, y, xi (geometry cartesian) or r, theta, xi
   (cylindrical).  Every xi position (gauss_center[2], pw_x[2], range, the
   perp_offset origin) is relative to the box start, as QPAD stores them.
   alpha != 0 tilts the transverse phase space (Twiss; twiss_beta = gamma
   sigma / uth as QPAD computes it); perp_offset adds sum_k c_k (xi -
   xi0)^k to x (index 0) or y (index 1).  Thermal momenta draw from
   Philox4x32-10 (counter = particle index) instead of gfortran's
   random_number. */
#define QPAD_PROF_UNIFORM 0
#define QPAD_PROF_GAUSSIAN 1
#define QPAD_PROF_PIECEWISE 2
#define QPAD_GEOM_CARTESIAN 0
#define QPAD_GEOM_CYLINDRICAL 1
#define QPAD_PW_MAX 32
#define QPAD_OFFSET_MAX 8
typedef struct {
    int32_t ppc[3];
    int32_t num_theta;
    int32_t geometry;
    int32_t profile[3];
    double qm, mass, density, gamma, den_min;
    double gauss_center[3], gauss_sigma[3];
    int32_t pw_n[3];
    double pw_x[3][QPAD_PW_MAX], pw_fx[3][QPAD_PW_MAX];
    double range_lo[3], range_hi[3];
    double uth[3];
    double alpha[2], twiss_ctr[2], twiss_beta[2];
    int32_t offset_n[2];
    double offset_xi0[2];
    double offset_c[2][QPAD_OFFSET_MAX];
    int32_t quiet;
    int32_t evol;
    uint64_t seed;
} qpad_beam_std_cfg;

typedef struct qpad_ctx qpad_ctx;

const char *qpad_last_error(void);
int qpad_abi_version(void);

/* simulation%alloc + %new without species/beams (simulation_class.f03:77-199,
   sim_fields_class.f03:91-160: fields, solver matrices) */
int qpad_create(const qpad_grid *grid, qpad_ctx **out);
/* simulation%del (simulation_class.f03:200-224) */
int qpad_destroy(qpad_ctx *ctx);
/* the grid the context was created with (options_class.f03:44-117 sizes
   and simulation settings), for host-side writers and drivers */
int qpad_get_grid(qpad_ctx *ctx, qpad_grid *out);

/* species2d%new: inject, initial charge, neutralising background
   (species2d_class.f03:75-134, part2d_class.f03:108-142,
   fdist2d_class.f03:289-357) */
int qpad_species_new(qpad_ctx *ctx, const qpad_species_cfg *cfg, int32_t *sp_id);
/* species2d%renew (species2d_class.f03:154-184) */
int qpad_species_renew(qpad_ctx *ctx, int32_t sp);
/* species2d%qdp: deposit rho - Jz of species sp and add it (plus the
   background) into q_spe (species2d_class.f03:186-208,
   part2d_class.f03:231-359).  The caller zeroes q_spe first, as QPAD does. */
int qpad_species_qdp(qpad_ctx *ctx, int32_t sp);
/* qpad_species_qdp with no device allocation: the deposit's per-cell partial
   rows live in a 2 MiB static device array instead of the species' own
   buffers; result bit for bit that of qpad_species_qdp.  No counterpart in
   QPAD: it is for callers that cannot allocate (code under stream capture,
   serving loops, harnesses that forbid hipMalloc).  Every species the
   library accepts fits (16 * nm * (nr + npmax / 1024 + 1) bytes); a larger
   one is refused with QPAD_E_INVALID.  The array is shared by every call on
   a GPU: calls that overlap on one GPU race. */
int qpad_species_qdp_noalloc(qpad_ctx *ctx, int32_t sp);
/* Per-species charge (species2d%q, its history written by species2d%cbq,
   species2d_class.f03:186-208,396-410), which QPAD's species charge_cyl_m
   diagnostic writes.  Off by default (it costs two small launches per slice
   and species and one history per species); qpad_species_charge_history(1)
   turns it on for every species, present and later.  With it on,
   qpad_slice / qpad_step3d keep it, qpad_species_qdp sets the species' q,
   qpad_species_cbq(sp, pos) adds the species' J_z and stores history slice
   pos, and qpad_get_species_charge reads history slice `slice`
   (1 x nm x (nr+2), as qpad_get_hist). */
int qpad_species_charge_history(qpad_ctx *ctx, int32_t enable);
int qpad_species_cbq(qpad_ctx *ctx, int32_t sp, int32_t pos);
int qpad_get_species_charge(qpad_ctx *ctx, int32_t sp, int32_t slice, double *out);
/* species2d%amjdp: deposit cu, acu (dJ/dxi part) and amu of species sp into
   the totals using fields e and b (species2d_class.f03:233-280;
   part2d_class.f03:746-1010 robust pusher, :478-744 standard pusher), with
   the 2D step dxi (the grid's dxi, or a sub-cycle's dxi / n) */
int qpad_species_amjdp(qpad_ctx *ctx, int32_t sp, double dxi);
/* species2d%push_u + %push_x (+ update_bound, move_part2d_comm):
   part2d_class.f03:1879-1965 (robust) or :1790-1877 (standard),
   :2221-2262, :2307-2353, with the 2D step dxi */
int qpad_species_push(qpad_ctx *ctx, int32_t sp, double dxi);
/* the sub-cycled push of species2d_subcyc (proj_subcyc/
   simulation_subcyc_class.f03:337-347): push_u, clamp_exp_fac
   (part2d_subcyc_class.f03:46-66: a particle whose gamma / (gamma - p_z)
   exceeds exp_fac_clamped (> 1) gets its momentum scaled so that the ratio
   equals it), push_x; standard and robust pushers */
int qpad_species_push_clamped(qpad_ctx *ctx, int32_t sp, double dxi, double exp_fac_clamped);
/* species2d%interp_psi (species2d_class.f03:439-452, part2d_class.f03:
   2264-2305): psi of the current slice at the particles of a standard-pusher
   species, which its amjdp and push read; no effect on a robust species.
   QPAD calls it after field_psi%solve (simulation_class.f03:357-359); the
   fused forms do it at the same point. */
int qpad_species_interp_psi(qpad_ctx *ctx, int32_t sp);

/* A non-uniform plasma (fdist2d's profiles, fdist2d_lib.f03:28-415): the
   density of the species' next injections (qpad_species_renew, the renew of
   qpad_step3d_end).  den_perp holds the transverse density get_den_perp
   gives at every lattice point, [num_theta][nr][ppc1][ppc2] in
   inject_fdist2d's loop order (fdist2d_class.f03:317-333: theta bin j,
   radial cell i, i1, i2; the point at rn = (i1 - 1/2)/ppc1 + i - 1,
   theta = ((i2 - 1/2)/ppc2 + j - 1) 2 pi / num_theta), den_lon the
   longitudinal factor get_den_lon gives at the injection's propagation
   distance s; a point with den_lon den_perp density < den_min is skipped,
   and a kept one carries the charge rn den_perp den_lon density coef.  The
   caller evaluates its profiles (at s = 0 before a renew of the new
   species, at s = i dt before the renew that ends 3D step i).  den_perp =
   NULL returns to the uniform plasma.  Not for neutrals. */
int qpad_species_set_profile(qpad_ctx *ctx, int32_t sp, const double *den_perp, double den_lon);

/* ---- sub-cycling (simulation.algorithm "subcycling", proj_subcyc) ----
   simulation_subcyc%new (simulation_subcyc_class.f03:76-88) with the deck's
   expansion_fac_max (> 0), expansion_fac_clamped (> 1) and dt_2d_min (in
   [dxi / 4096, dxi]); on = 0 returns to QPAD's standard slice.  Each fused
   slice then takes the largest gamma / (gamma - p_z) of any particle
   (get_exp_fac_max), runs ceiling(factor / expansion_fac_max) sub-cycles
   of dxi / n (no more than floor(dxi / dt_2d_min); get_subcyc_step,
   :431-449), each a charge deposit, the solves, the predictor-corrector
   iteration with step dxi / n and the clamped push (neutrals ionised with
   their own step dxi in each), and closes the slice once (:230-390).
   qpad_slice's iteration count is the sum over the sub-cycles.  Not with a
   laser, the PGC pushers, the seed network or a batch. */
int qpad_set_subcycling(qpad_ctx *ctx, int32_t on, double exp_fac_max, double exp_fac_clamped, double dt_2d_min);
/* sim_plasma_subcyc%get_exp_fac_max (sim_plasma_subcyc_class.f03:58-86,
   part2d_subcyc_class.f03:27-44): the largest gamma / (gamma - p_z) over
   every particle of every species and neutral, at least 1 */
int qpad_exp_fac_max(qpad_ctx *ctx, double *out);
/* the last fused slice's sub-cycles n, their 2D step and the factor they
   came from (1, dxi, 0 before any sub-cycled slice) */
int qpad_last_subcycles(qpad_ctx *ctx, int32_t *n, double *dxi_sub, double *exp_fac);

/* ---- neutral gas and field ionisation ----
   neutrals(i) section (sim_plasma_class.f03:148-176, neutral_class.f03:
   408-574): a uniform gas of `element` (atomic number 1, 2, 3, 6, 7, 8, 10
   or 18) followed up to ion_max ionisation levels, ppc and num_theta as a
   species (the release lattice), q and m of its electrons, push type
   QPAD_PUSH2_STD or QPAD_PUSH2_ROBUST; n0 (simulation.n0, cm^-3) gives the
   plasma frequency of the ADK rates. */
typedef struct {
    int32_t ppc1, ppc2, num_theta;
    double qm, mass, density, den_min;
    int32_t element, ion_max, push_type;
    double n0;
} qpad_neutral_cfg;
/* neutral%new: a species whose particles are the electrons the field
   releases (none yet; qpad_species_* calls take its id, qpad_get_species_qn
   its ion charge rho_ion); the fused slice ionises it with the slice's
   final E and deposits the ions at the next slice's start */
int qpad_neutral_new(qpad_ctx *ctx, const qpad_neutral_cfg *cfg, int32_t *sp_id);
/* neutral%ion_deposit(q_spe) (neutral_class.f03:904-930): rho_ion += the ion
   charge of the electrons the last update released; q_spe += rho_ion */
int qpad_neutral_ion_deposit(qpad_ctx *ctx, int32_t sp);
/* neutral%update(e, psi, s) (neutral_class.f03:576-598, 600-837): the level
   fractions from E (QPAD_F_E), and the released electrons appended at rest */
int qpad_neutral_update(qpad_ctx *ctx, int32_t sp);
/* the level fractions [num_theta][nr][ion_max + 2] (levels 1..ion_max, the
   neutral residue, the discrete total ion charge; neutral%multi_ion) */
int qpad_get_neutral_levels(qpad_ctx *ctx, int32_t sp, double *out);
/* rho_ion of history slice s (neutral%cbq, neutral_class.f03:1158-1176) */
int qpad_get_ion_density(qpad_ctx *ctx, int32_t sp, int32_t s, double *out);
/* the ADK coefficients (C1, C2, C3) of a level (neutral_class.f03:23-299) */
int qpad_neutral_adk(int32_t element, int32_t level, double *out);

/* ---- laser (PGC pushers) ----
   The laser envelope the PGC pushers push in: the real and imaginary part
   of the complex envelope a of every slice, [(nxi + 3)][nm][nr + 2] each,
   slices -1 .. nxi + 1 (QPAD's field_laser cfr, cfi f2, field_laser_class.f03:
   104-180, with two guard slices below the box and one above,
   sim_lasers_class.f03:107-108; with several lasers their sum).  Replaces
   the envelope QPAD's laser solver keeps (sim_lasers%advance,
   sim_lasers_class.f03:202-223, which stays with the caller between 3D
   steps).  The first call (or the first PGC species)
   switches the context's slices to the laser path: every slice then takes
   its envelope and gradients (qpad_laser_slice) and deposits chi
   (qpad_deposit_chi) into the QPAD_F_CHI history the laser solver reads. */
int qpad_set_laser(qpad_ctx *ctx, const double *ar_hist, const double *ai_hist);
/* laser_all of slice j: zero, copy_slice(j), set_grad(j), gather
   (simulation_class.f03:362-367, field_laser_class.f03:637-750, :929-1012):
   fields QPAD_F_AR, QPAD_F_AI, QPAD_F_AR_GRAD, QPAD_F_AI_GRAD */
int qpad_laser_slice(qpad_ctx *ctx, int32_t j);
/* sim_lasers%deposit_chi(spe, j) (sim_lasers_class.f03:176-200,
   part2d_class.f03:361-476): chi of every species into QPAD_F_CHI and its
   history slice j */
int qpad_deposit_chi(qpad_ctx *ctx, int32_t j);

/* field_psi%solve(q_spe) (field_psi_class.f03:218-254) */
int qpad_field_solve_psi(qpad_ctx *ctx);
/* field_e%solve(jay): Ez of field `dst` from current `jay`
   (field_e_class.f03:148-334) */
int qpad_field_solve_ez(qpad_ctx *ctx, int32_t dst, int32_t jay);
/* field_e%solve(b, psi): E_perp of `dst` from `b` and psi
   (field_e_class.f03:412-514) */
int qpad_field_solve_et(qpad_ctx *ctx, int32_t dst, int32_t b);
/* field_b%solve(jay): Bz of `dst` from `jay` (field_b_class.f03:760-796) */
int qpad_field_solve_bz(qpad_ctx *ctx, int32_t dst, int32_t jay);
/* field_b%solve(rho): beam B_perp of slice j from the q_beam history
   (field_b_class.f03:798-834, with q_beam%copy_slice(j, p_copy_2to1)) */
int qpad_field_solve_bt_beam(qpad_ctx *ctx, int32_t j);
/* field_djdxi%solve(acu, amu) (field_src_class.f03:273-405) */
int qpad_field_solve_djdxi(qpad_ctx *ctx);
/* field_b%solve(djdxi, jay): plasma B_perp predictor-corrector solve into
   b_spe (field_b_class.f03:836-879) */
int qpad_field_solve_bt_iter(qpad_ctx *ctx);
/* add_f1(b_spe, b_beam, b) (ufield_class.f03:842-904, called at simulation_class.f03:375,413) */
int qpad_field_add_b(qpad_ctx *ctx);
/* convergence_tester(b_spe, 2, op) (simulation_class.f03:522-606);
   op 0 = 'record', 1 = 'compare' (rel/abs returned for 'compare') */
int qpad_convergence_tester(qpad_ctx *ctx, int32_t op, double *rel_res, double *abs_res);

/* beam3d%qdp for all beams into the q_beam history (beam3d_class.f03:193-221,
   part3d_class.f03:221-356) */
int qpad_beam_qdp(qpad_ctx *ctx);
/* beam3d%push (reduced pusher) for all beams, with update_bound
   (beam3d_class.f03:223-254, part3d_class.f03:477-576, :640-790) */
int qpad_beam_push(qpad_ctx *ctx);
/* beam3d%new with a random Gaussian profile (beam3d_class.f03:121-177,
   fdist3d_rnd_class.f03:360-513); draws from Philox4x32-10 */
int qpad_beam_new(qpad_ctx *ctx, const qpad_beam_cfg *cfg, int32_t *beam_id);
/* beam3d%new with the lattice profile ("standard", every shipped QPAD deck):
   beam3d_class.f03:121-177, fdist3d_std_class.f03:439-613; the beam holds
   exactly the injected particles (QPAD's npmax is not needed) */
int qpad_beam_new_std(qpad_ctx *ctx, const qpad_beam_std_cfg *cfg, int32_t *beam_id);
/* The lattice beam with a density the caller evaluates: the profiles of
   fdist3d_std_lib.f03 this library does not evaluate itself (super-gaussian
   :129-172, parabolic :174-214, rational :216-286, analytic :46-84 with its
   math_func of x, y and z) or any other.  den holds the density of every
   lattice point of the box xi cells k0 .. k0 + nk - 1 (1-based), theta
   sectors 1 .. num_theta, r cells 1 .. ni, in the order k > j > i > i3 > i2
   > i1 of inject_fdist3d_std (fdist3d_std_class.f03:439-530; the point at
   rn = (i1 - 1/2)/ppc1 + i - 1, theta = ((i2 - 1/2)/ppc2 + j - 1) 2 pi /
   num_theta, zn = (i3 - 1/2)/ppc3 + k - 1, in cfg's geometry), without the
   factor cfg->density; the box must hold the one qpad_beam_lattice_box
   gives for cfg's ranges.  cfg's profile fields are not read; everything
   else (ranges, den_min, quiet start, Twiss, offsets, momenta) is as in
   qpad_beam_new_std. */
int qpad_beam_lattice_box(qpad_ctx *ctx, const qpad_beam_std_cfg *cfg, int32_t *k0, int32_t *nk, int32_t *ni);
int qpad_beam_new_std_density(qpad_ctx *ctx, const qpad_beam_std_cfg *cfg, const double *den, int32_t k0,
                              int32_t nk, int32_t ni, int32_t *beam_id);
/* beam3d%new with given particles: the "file" profile
   (fdist3d_file_class.f03:154-343), whose host side reads and places them
   (qpad_diag_read_beam_file); x3 relative to the box start */
int qpad_beam_new_arrays(qpad_ctx *ctx, double qm, double mass, int32_t evol, int64_t n, const double *x1,
                         const double *x2, const double *x3, const double *p1, const double *p2, const double *p3,
                         const double *q, int32_t *beam_id);

/* The fused forms of simulation%run (simulation_class.f03:294-512):
   qpad_step3d_begin = beam deposit, beam B for every slice, field zeroing
   (:297-340); qpad_slice = one xi-slice (:342-469) with every predictor-
   corrector iteration (iters_out = iterations executed); qpad_step3d_end =
   beam push and species renew (:485-501); qpad_step3d = all three. */
int qpad_step3d_begin(qpad_ctx *ctx);
int qpad_slice(qpad_ctx *ctx, int32_t j, int32_t *iters_out);
int qpad_step3d_end(qpad_ctx *ctx);
int qpad_step3d(qpad_ctx *ctx, int32_t *total_iters_out);
/* iterations of the last 3D step of qpad_step3d or qpad_batch_step3d:
   summed over its slices, and the number of its slices that ended at
   iter_max without meeting the tolerance (no counterpart in QPAD, which
   logs neither) */
int qpad_step_iterations(qpad_ctx *ctx, int32_t *total_iters, int32_t *slices_at_iter_max);
/* How qpad_slice / qpad_step3d queue the predictor-corrector iterations (no
   counterpart in QPAD, whose loop is serial on the host; results do not
   depend on it).  adaptive = 0: all iter_max iterations are queued at once
   and the ones after convergence return immediately on the device.
   adaptive != 0 (default, margin 0): the slice queues as many iterations as
   the previous slice needed plus `margin`, then one further iteration at a
   time while the device reports no convergence (one host wait per further
   iteration, overlapped with that iteration's field solves). */
int qpad_iteration_policy(qpad_ctx *ctx, int32_t adaptive, int32_t margin);
/* Launch shapes of the dominant kernels (no counterpart in QPAD):
   dep_cells = radial cells, one wave each, per workgroup of the current and
   charge deposits (2, 4 or 8; default 4; results do not depend on it);
   dep_mfma = 1: the current deposit sums each cell's particles on the matrix
   cores (FP64 MFMA), 0: with the wave reduce-scatter; the two forms add the
   same terms in different orders (results equal to rounding), so the form
   is chosen here and never by the autotuner (default: 1 for max_mode >= 3,
   where it measured faster on MI355X, else 0). */
typedef struct {
    int32_t dep_cells;
    int32_t dep_mfma;
} qpad_launch_shape;
int qpad_get_launch_shape(qpad_ctx *ctx, qpad_launch_shape *out);
int qpad_set_launch_shape(qpad_ctx *ctx, const qpad_launch_shape *shape);
/* Autotuner: times every candidate dep_cells on the context's current
   particles and fields in the chosen deposit form (reps launches each; the
   deposits only rewrite scratch partials, so no result changes) and keeps
   the fastest.  ms_out: mean ms per candidate dep_cells = 2, 4, 8
   (3 entries) or NULL. */
int qpad_tune_launch(qpad_ctx *ctx, int32_t reps, double *ms_out);

/* state exchange with the host (diagnostics, restart, parity checks) */
int qpad_get_field(qpad_ctx *ctx, int32_t fid, double *out);
int qpad_set_field(qpad_ctx *ctx, int32_t fid, const double *in);
/* history slice s (0..nxi+1) of psi, e, b, b_spe, e_spe, cu, q_spe, q_beam
   (and b_beam, whose third component is always zero) */
int qpad_get_hist(qpad_ctx *ctx, int32_t fid, int32_t s, double *out);
int qpad_set_hist(qpad_ctx *ctx, int32_t fid, int32_t s, const double *in);
int qpad_species_count(qpad_ctx *ctx, int32_t sp, int64_t *npp);
/* any pointer may be NULL to skip that array; arrays hold npp entries in the
   library's (radial-cell sorted) order */
int qpad_get_species(qpad_ctx *ctx, int32_t sp, double *x1, double *x2, double *p1,
                     double *p2, double *p3, double *q, double *gamma, double *psi,
                     int32_t *id);
/* replace species sp's particles (n entries; gamma/psi/id may be NULL) */
int qpad_set_species(qpad_ctx *ctx, int32_t sp, int64_t n, const double *x1,
                     const double *x2, const double *p1, const double *p2,
                     const double *p3, const double *q, const double *gamma,
                     const double *psi, const int32_t *id);
/* the species' neutralising background qn (1 x nm x (nr+2)) */
int qpad_get_species_qn(qpad_ctx *ctx, int32_t sp, double *out);
int qpad_set_species_qn(qpad_ctx *ctx, int32_t sp, const double *in);
int qpad_beam_count(qpad_ctx *ctx, int32_t ib, int64_t *npp);
int qpad_get_beam(qpad_ctx *ctx, int32_t ib, double *x1, double *x2, double *x3,
                  double *p1, double *p2, double *p3, double *q);
int qpad_set_beam(qpad_ctx *ctx, int32_t ib, int64_t n, const double *x1,
                  const double *x2, const double *x3, const double *p1,
                  const double *p2, const double *p3, const double *q);

/* timing of the dominant kernels (HIP events on the context's stream),
   sampled on the first predictor-corrector iteration of every slice (which
   always does work; events around every launch would slow the step):
   which = 0: plasma current deposit (amjdeposit), 1: plasma push (every
   slice), 2: the field sources and solves of the first iteration.  Returns
   the summed milliseconds and the number of timed launches since the last
   reset. */
int qpad_kernel_time(qpad_ctx *ctx, int32_t which, double *ms, int64_t *launches);
int qpad_kernel_time_reset(qpad_ctx *ctx, int32_t enable);

/* ---- learned B_perp seed of the predictor-corrector iteration ----
   No counterpart in QPAD.  QPAD seeds the iteration of slice xi with the
   converged B_perp of slice xi-1 (b_spe is zeroed only at the 3D-step start,
   simulation_class.f03:326, and read by the first 'record', :370-375).  The
   seed network replaces that seed by 2 B_perp(xi-1) - B_perp(xi-2) + dB: the
   first-order extrapolation in xi plus a correction dB predicted by a 1D
   convolution network over r from the fields slice xi-1 left (its B_perp
   and charge, the J predictor of slice xi) and the beam charge of slice xi
   (the output layer starts at zero: untrained, the seed is the
   extrapolation).  The iteration still runs to the same tolerance and
   iter_max, so the plain method still decides convergence.  Since the inputs are
   ready when slice xi-1 ends, qpad_step3d computes each seed on a second
   stream beside the previous slice's push and sort.

   Network: `layers` conv1d layers over the radial nodes 1..nr, kernel 3,
   zero padding, SiLU between layers, `width` hidden channels.  Inputs per
   node (8 (2M+1) + 1 channels, each group scaled by its maximum over the
   slice):
     B_r, B_phi of b_spe (slice xi-1);  B_r, B_phi of b_spe(xi-1) -
     b_spe(xi-2);  J_r, J_phi of cu (the predictor of slice xi);  q_spe of
     slice xi-1;  q_beam of slice xi;  r / r_max.
   Output: dB_r, dB_phi (2 (2M+1) channels) in units of out_scale times the
   B_perp scale; training target: (converged - extrapolated) in the same
   units, clipped to [-1, 1] (so the correction never moves the seed by more
   than out_scale max|B_perp| from the extrapolation).
   Weights and activations are BF16, products accumulate in FP32 (matrix
   cores); master weights and Adam moments are FP32.

   Parameter layout (FP32, qpad_seednet_get_params): layer after layer,
   W[3][cout_p][cin_p] then bias[cout_p], with the padded channel counts of
   qpad_seednet_info; padding entries are zero and stay zero. */
typedef struct {
    int32_t width;   /* hidden channels (64) */
    int32_t layers;  /* conv layers, 2..8 (4) */
    float lr;        /* Adam step size (1e-3) */
    float beta1;     /* (0.9) */
    float beta2;     /* (0.999) */
    float eps;       /* (1e-8) */
    float out_scale; /* correction units, fraction of max|B_perp| (0.02) */
    uint64_t seed;   /* weight initialisation */
} qpad_seednet_cfg;

typedef struct {
    int32_t layers;
    int32_t cin;          /* input channels used: 8 (2M+1) + 1 */
    int32_t cout;         /* output channels used: 2 (2M+1) */
    int32_t npos;         /* radial nodes (nr) */
    int32_t cin_p[8];     /* padded input channels of each layer */
    int32_t cout_p[8];    /* padded output channels of each layer */
    int64_t nparams;      /* FP32 parameters, padding included */
    int64_t steps;        /* Adam steps taken */
} qpad_seednet_info;

/* create (or re-create) the context's seed network; both modes start off */
int qpad_seednet_create(qpad_ctx *ctx, const qpad_seednet_cfg *cfg);
int qpad_seednet_get_info(qpad_ctx *ctx, qpad_seednet_info *out);
/* seed != 0: qpad_slice / qpad_step3d seed each slice's iteration with the
   prediction; train != 0: they take one Adam step per slice on (inputs,
   converged B_perp - extrapolation) once the slice has converged */
int qpad_seednet_mode(qpad_ctx *ctx, int32_t seed, int32_t train);
int qpad_seednet_get_params(qpad_ctx *ctx, float *theta, int64_t n);
int qpad_seednet_set_params(qpad_ctx *ctx, const float *theta, int64_t n);
/* training losses of Adam steps first .. first + n - 1 (the last 65536 are
   kept; mean squared error of the normalised output) */
int qpad_seednet_losses(qpad_ctx *ctx, int64_t first, int64_t n, float *out);
/* Fine-grained form for a driver that calls QPAD's procedures one by one:
   qpad_seednet_seed(j) first in slice j, with the fields slice j - 1 left
   (before q_spe is recomputed; scales and packs the inputs, and in seed
   mode sets b_spe's B_perp to the seed), qpad_seednet_learn(j) once slice
   j has converged (one Adam step; needs qpad_seednet_seed of the same slice
   first). */
int qpad_seednet_seed(qpad_ctx *ctx, int32_t j);
int qpad_seednet_learn(qpad_ctx *ctx, int32_t j);
/* The network on host inputs (tests, offline use): x [npos][cin] FP32
   (rounded to BF16 on the device), t [npos][cout] FP32 target or NULL.
   Returns y [npos][cout] and, with t, the loss and the gradient of every
   parameter (grad: nparams entries or NULL); update != 0 also takes the
   Adam step.  npos <= nr.  Slice state is not touched. */
int qpad_seednet_eval(qpad_ctx *ctx, const float *x, const float *t, int32_t npos,
                      float *y, float *loss, float *grad, int32_t update);

/* ---- batched instances ----
   No counterpart in QPAD, whose runs are one simulation each (parameter
   scans run as separate jobs).  A batch steps several independent contexts
   (different nr, modes, beams) in one call, each context on its own
   stream, so the kernels of different instances run side by side on the
   device.  Every context's results equal its own qpad_step3d.  The
   contexts stay usable on their own. */
typedef struct qpad_batch qpad_batch;
int qpad_batch_create(qpad_ctx *const *ctxs, int32_t n, qpad_batch **out);
int qpad_batch_destroy(qpad_batch *batch);
/* mode 0 (default): eager, one host thread per context queues its slices
   with the context's own iteration policy (qpad_iteration_policy);
   mode 1: each context's slice loop is captured once into a hipGraph (all
   iter_max iterations queued; the ones after convergence return
   immediately on the device) and replayed every step; needs an even nxi
   (the particle buffers swap once per slice);
   mode 2: eager from the calling thread alone, which polls every context's
   pending iteration decision and serves the ready ones. */
int qpad_batch_mode(qpad_batch *batch, int32_t mode);
/* one 3D step of every context; total_iters_out: n entries or NULL */
int qpad_batch_step3d(qpad_batch *batch, int32_t *total_iters_out);
/* qpad_tune_launch on every context (captured graphs are dropped) */
int qpad_batch_tune(qpad_batch *batch, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif
