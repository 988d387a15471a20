! qpad_hip_mod.f03 -- ISO_C_BINDING interface to libqpad_hip.so for QPAD's
! Fortran side.  One interface per entry point of include/qpad_hip.h; the
! derived types mirror the C structs field for field (bind(C)).
!
! QPAD never checks return codes: its procedures stop through write_err
! (source/sysutil_module.f03:165-183).  qpad_hip_check turns a nonzero
! return into that same path: the caller passes its own error routine.
module qpad_hip_mod

  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: QPAD_HIP_ABI_VERSION = 2

  integer(c_int), parameter, public :: QPAD_OK = 0, QPAD_E_INVALID = -1, &
    QPAD_E_HIP = -2, QPAD_E_NOMEM = -3, QPAD_E_STATE = -4

  integer(c_int), parameter, public :: QPAD_F_PSI = 0, QPAD_F_E = 1, &
    QPAD_F_E_SPE = 2, QPAD_F_B = 3, QPAD_F_B_SPE = 4, QPAD_F_B_BEAM = 5, &
    QPAD_F_CU = 6, QPAD_F_ACU = 7, QPAD_F_AMU = 8, QPAD_F_DCU = 9, &
    QPAD_F_Q_SPE = 10, QPAD_F_Q_BEAM = 11, QPAD_F_AR = 12, QPAD_F_AI = 13, &
    QPAD_F_AR_GRAD = 14, QPAD_F_AI_GRAD = 15, QPAD_F_CHI = 16

  integer(c_int), parameter, public :: QPAD_BND_ZERO = 2, QPAD_BND_OPEN = 3
  integer(c_int), parameter, public :: QPAD_PUSH2_STD = 0, QPAD_PUSH2_ROBUST = 1, &
    QPAD_PUSH2_STD_PGC = 4, QPAD_PUSH2_ROBUST_PGC = 5

  type, bind(C), public :: qpad_grid
    integer(c_int32_t) :: nr, nxi, max_mode
    real(c_double) :: dr, dxi
    real(c_double) :: relax_fac
    integer(c_int32_t) :: bnd
    integer(c_int32_t) :: iter_max
    real(c_double) :: iter_reltol, iter_abstol
    real(c_double) :: dt
  end type qpad_grid

  type, bind(C), public :: qpad_species_cfg
    integer(c_int32_t) :: ppc1, ppc2, num_theta
    real(c_double) :: qm, mass, density
    real(c_double) :: uth(3)
    integer(c_int32_t) :: neutralized
    real(c_double) :: den_min
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: push_type
  end type qpad_species_cfg

  ! neutrals(i) of the input deck: element (atomic number) and ion_max set
  ! the ADK levels followed, n0 the plasma frequency of the ADK rates
  type, bind(C), public :: qpad_neutral_cfg
    integer(c_int32_t) :: ppc1, ppc2, num_theta
    real(c_double) :: qm, mass, density, den_min
    integer(c_int32_t) :: element, ion_max, push_type
    real(c_double) :: n0
  end type qpad_neutral_cfg

  integer(c_int), parameter, public :: QPAD_PROF_UNIFORM = 0, QPAD_PROF_GAUSSIAN = 1, &
    QPAD_PROF_PIECEWISE = 2, QPAD_GEOM_CARTESIAN = 0, QPAD_GEOM_CYLINDRICAL = 1, &
    QPAD_PW_MAX = 32, QPAD_OFFSET_MAX = 8

  ! lattice beam (profile_type "standard"); C arrays [3][32] are Fortran
  ! (32, 3), so pw_x(:, d) is direction d
  type, bind(C), public :: qpad_beam_std_cfg
    integer(c_int32_t) :: ppc(3), num_theta, geometry, profile(3)
    real(c_double) :: qm, mass, density, gamma, den_min
    real(c_double) :: gauss_center(3), gauss_sigma(3)
    integer(c_int32_t) :: pw_n(3)
    real(c_double) :: pw_x(32, 3), pw_fx(32, 3)
    real(c_double) :: range_lo(3), range_hi(3)
    real(c_double) :: uth(3)
    real(c_double) :: alpha(2), twiss_ctr(2), twiss_beta(2)
    integer(c_int32_t) :: offset_n(2)
    real(c_double) :: offset_xi0(2)
    real(c_double) :: offset_c(8, 2)
    integer(c_int32_t) :: quiet, evol
    integer(c_int64_t) :: seed
  end type qpad_beam_std_cfg

  type, bind(C), public :: qpad_beam_cfg
    integer(c_int64_t) :: total_num
    real(c_double) :: total_charge
    real(c_double) :: qm, mass, gamma
    real(c_double) :: center(3), sigma(3)
    real(c_double) :: range_lo(3), range_hi(3)
    real(c_double) :: uth(3)
    integer(c_int32_t) :: quiet
    integer(c_int32_t) :: evol
    integer(c_int64_t) :: seed
  end type qpad_beam_cfg

  ! learned B_perp seed of the iteration (no counterpart in QPAD)
  type, bind(C), public :: qpad_seednet_cfg
    integer(c_int32_t) :: width, layers
    real(c_float) :: lr, beta1, beta2, eps, out_scale
    integer(c_int64_t) :: seed
  end type qpad_seednet_cfg

  type, bind(C), public :: qpad_seednet_info
    integer(c_int32_t) :: layers, cin, cout, npos
    integer(c_int32_t) :: cin_p(8), cout_p(8)
    integer(c_int64_t) :: nparams, steps
  end type qpad_seednet_info

  type, bind(C), public :: qpad_launch_shape
    integer(c_int32_t) :: dep_cells, dep_mfma
  end type qpad_launch_shape

  public :: qpad_hip_check, qpad_error_message

  interface

    function qpad_last_error() bind(C, name="qpad_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function

    function qpad_abi_version() bind(C, name="qpad_abi_version") result(v)
      import :: c_int
      integer(c_int) :: v
    end function

    ! simulation%alloc/%new without species and beams
    function qpad_create(grid, ctx) bind(C, name="qpad_create") result(rc)
      import :: c_int, c_ptr, qpad_grid
      type(qpad_grid), intent(in) :: grid
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function

    function qpad_destroy(ctx) bind(C, name="qpad_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_get_grid(ctx, grid) bind(C, name="qpad_get_grid") result(rc)
      import :: c_int, c_ptr, qpad_grid
      type(c_ptr), value :: ctx
      type(qpad_grid), intent(out) :: grid
      integer(c_int) :: rc
    end function

    ! species2d%new / %renew / %qdp / %amjdp / %push_u + %push_x
    function qpad_species_new(ctx, cfg, sp_id) bind(C, name="qpad_species_new") result(rc)
      import :: c_int, c_int32_t, c_ptr, qpad_species_cfg
      type(c_ptr), value :: ctx
      type(qpad_species_cfg), intent(in) :: cfg
      integer(c_int32_t), intent(out) :: sp_id
      integer(c_int) :: rc
    end function

    function qpad_species_renew(ctx, sp) bind(C, name="qpad_species_renew") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int) :: rc
    end function

    function qpad_species_qdp(ctx, sp) bind(C, name="qpad_species_qdp") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int) :: rc
    end function

    ! the same with the deposit's partial rows in a static device array (no
    ! device allocation; calls that overlap on one GPU race)
    function qpad_species_qdp_noalloc(ctx, sp) bind(C, name="qpad_species_qdp_noalloc") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int) :: rc
    end function

    ! species2d%q history (cbq) for the species charge_cyl_m diagnostic
    function qpad_species_charge_history(ctx, enable) bind(C, name="qpad_species_charge_history") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
      integer(c_int) :: rc
    end function

    function qpad_species_cbq(ctx, sp, pos) bind(C, name="qpad_species_cbq") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp, pos
      integer(c_int) :: rc
    end function

    function qpad_get_species_charge(ctx, sp, slice, out) bind(C, name="qpad_get_species_charge") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp, slice
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_species_amjdp(ctx, sp, dxi) bind(C, name="qpad_species_amjdp") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      real(c_double), value :: dxi
      integer(c_int) :: rc
    end function

    function qpad_species_push(ctx, sp, dxi) bind(C, name="qpad_species_push") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      real(c_double), value :: dxi
      integer(c_int) :: rc
    end function

    ! the lattice beam with a density the caller evaluates over the box of its ranges
    function qpad_beam_lattice_box(ctx, cfg, k0, nk, ni) bind(C, name="qpad_beam_lattice_box") result(rc)
      import :: c_int, c_int32_t, c_ptr, qpad_beam_std_cfg
      type(c_ptr), value :: ctx
      type(qpad_beam_std_cfg), intent(in) :: cfg
      integer(c_int32_t), intent(out) :: k0, nk, ni
      integer(c_int) :: rc
    end function

    function qpad_beam_new_std_density(ctx, cfg, den, k0, nk, ni, beam_id) &
        bind(C, name="qpad_beam_new_std_density") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double, qpad_beam_std_cfg
      type(c_ptr), value :: ctx
      type(qpad_beam_std_cfg), intent(in) :: cfg
      real(c_double), intent(in) :: den(*)
      integer(c_int32_t), value :: k0, nk, ni
      integer(c_int32_t), intent(out) :: beam_id
      integer(c_int) :: rc
    end function

    ! a non-uniform plasma: the transverse density of every lattice point and
    ! the longitudinal factor of the next injections
    function qpad_species_set_profile(ctx, sp, den_perp, den_lon) bind(C, name="qpad_species_set_profile") &
        result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      type(c_ptr), value :: den_perp
      real(c_double), value :: den_lon
      integer(c_int) :: rc
    end function

    ! species2d_subcyc's push: push_u, clamp_exp_fac, push_x
    function qpad_species_push_clamped(ctx, sp, dxi, exp_fac_clamped) bind(C, name="qpad_species_push_clamped") &
        result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      real(c_double), value :: dxi, exp_fac_clamped
      integer(c_int) :: rc
    end function

    ! simulation_subcyc%new, get_exp_fac_max, the last slice's sub-cycles
    function qpad_set_subcycling(ctx, on, exp_fac_max, exp_fac_clamped, dt_2d_min) &
        bind(C, name="qpad_set_subcycling") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
      real(c_double), value :: exp_fac_max, exp_fac_clamped, dt_2d_min
      integer(c_int) :: rc
    end function

    function qpad_exp_fac_max(ctx, out) bind(C, name="qpad_exp_fac_max") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out
      integer(c_int) :: rc
    end function

    function qpad_last_subcycles(ctx, n, dxi_sub, exp_fac) bind(C, name="qpad_last_subcycles") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: n
      real(c_double), intent(out) :: dxi_sub, exp_fac
      integer(c_int) :: rc
    end function

    ! iterations of the last 3D step: total, and slices that ended at iter_max
    function qpad_step_iterations(ctx, total_iters, slices_at_iter_max) &
        bind(C, name="qpad_step_iterations") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: total_iters, slices_at_iter_max
      integer(c_int) :: rc
    end function

    ! species2d%interp_psi (standard pusher; no effect on a robust species)
    function qpad_species_interp_psi(ctx, sp) bind(C, name="qpad_species_interp_psi") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int) :: rc
    end function

    ! the laser envelope history (field_laser cfr, cfi f2), each
    ! ar(0:nr+1, nm, -1:nxi+1) in Fortran order
    function qpad_set_laser(ctx, ar, ai) bind(C, name="qpad_set_laser") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ar(*), ai(*)
      integer(c_int) :: rc
    end function

    function qpad_laser_slice(ctx, j) bind(C, name="qpad_laser_slice") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: j
      integer(c_int) :: rc
    end function

    function qpad_deposit_chi(ctx, j) bind(C, name="qpad_deposit_chi") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: j
      integer(c_int) :: rc
    end function

    ! field_psi%solve, field_e%solve(jay) / (b, psi), field_b%solve(jay) /
    ! (rho) / (djdxi, jay), field_djdxi%solve, add_f1, convergence_tester
    function qpad_field_solve_psi(ctx) bind(C, name="qpad_field_solve_psi") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_field_solve_ez(ctx, dst, jay) bind(C, name="qpad_field_solve_ez") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dst, jay
      integer(c_int) :: rc
    end function

    function qpad_field_solve_et(ctx, dst, b) bind(C, name="qpad_field_solve_et") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dst, b
      integer(c_int) :: rc
    end function

    function qpad_field_solve_bz(ctx, dst, jay) bind(C, name="qpad_field_solve_bz") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dst, jay
      integer(c_int) :: rc
    end function

    function qpad_field_solve_bt_beam(ctx, j) bind(C, name="qpad_field_solve_bt_beam") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: j
      integer(c_int) :: rc
    end function

    function qpad_field_solve_djdxi(ctx) bind(C, name="qpad_field_solve_djdxi") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_field_solve_bt_iter(ctx) bind(C, name="qpad_field_solve_bt_iter") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_field_add_b(ctx) bind(C, name="qpad_field_add_b") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_convergence_tester(ctx, op, rel_res, abs_res) &
        bind(C, name="qpad_convergence_tester") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: op
      real(c_double), intent(out) :: rel_res, abs_res
      integer(c_int) :: rc
    end function

    ! beam3d%qdp / %push / %new
    function qpad_beam_qdp(ctx) bind(C, name="qpad_beam_qdp") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_beam_push(ctx) bind(C, name="qpad_beam_push") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_beam_new(ctx, cfg, beam_id) bind(C, name="qpad_beam_new") result(rc)
      import :: c_int, c_int32_t, c_ptr, qpad_beam_cfg
      type(c_ptr), value :: ctx
      type(qpad_beam_cfg), intent(in) :: cfg
      integer(c_int32_t), intent(out) :: beam_id
      integer(c_int) :: rc
    end function

    ! beam3d%new with given particles (the "file" profile)
    function qpad_beam_new_arrays(ctx, qm, mass, evol, n, x1, x2, x3, p1, p2, p3, q, beam_id) &
        bind(C, name="qpad_beam_new_arrays") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: qm, mass
      integer(c_int32_t), value :: evol
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x1(*), x2(*), x3(*), p1(*), p2(*), p3(*), q(*)
      integer(c_int32_t), intent(out) :: beam_id
      integer(c_int) :: rc
    end function

    ! beam3d%new with the lattice ("standard") profile
    function qpad_beam_new_std(ctx, cfg, beam_id) bind(C, name="qpad_beam_new_std") result(rc)
      import :: c_int, c_int32_t, c_ptr, qpad_beam_std_cfg
      type(c_ptr), value :: ctx
      type(qpad_beam_std_cfg), intent(in) :: cfg
      integer(c_int32_t), intent(out) :: beam_id
      integer(c_int) :: rc
    end function

    ! fused simulation%run pieces
    function qpad_step3d_begin(ctx) bind(C, name="qpad_step3d_begin") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_slice(ctx, j, iters) bind(C, name="qpad_slice") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: j
      integer(c_int32_t), intent(out) :: iters
      integer(c_int) :: rc
    end function

    function qpad_step3d_end(ctx) bind(C, name="qpad_step3d_end") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function qpad_step3d(ctx, total_iters) bind(C, name="qpad_step3d") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: total_iters
      integer(c_int) :: rc
    end function

    ! host <-> device state: fields as (0:nr+1, 2*max_mode+1, ncomp)
    function qpad_get_field(ctx, fid, out) bind(C, name="qpad_get_field") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: fid
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_set_field(ctx, fid, in) bind(C, name="qpad_set_field") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: fid
      real(c_double), intent(in) :: in(*)
      integer(c_int) :: rc
    end function

    function qpad_get_hist(ctx, fid, s, out) bind(C, name="qpad_get_hist") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: fid, s
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_set_hist(ctx, fid, s, in) bind(C, name="qpad_set_hist") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: fid, s
      real(c_double), intent(in) :: in(*)
      integer(c_int) :: rc
    end function

    function qpad_species_count(ctx, sp, npp) bind(C, name="qpad_species_count") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int64_t), intent(out) :: npp
      integer(c_int) :: rc
    end function

    ! pass c_null_ptr (via c_loc of nothing) for arrays to skip
    function qpad_get_species(ctx, sp, x1, x2, p1, p2, p3, q, gamma, psi, id) &
        bind(C, name="qpad_get_species") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      type(c_ptr), value :: x1, x2, p1, p2, p3, q, gamma, psi, id
      integer(c_int) :: rc
    end function

    function qpad_set_species(ctx, sp, n, x1, x2, p1, p2, p3, q, gamma, psi, id) &
        bind(C, name="qpad_set_species") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int64_t), value :: n
      type(c_ptr), value :: x1, x2, p1, p2, p3, q, gamma, psi, id
      integer(c_int) :: rc
    end function

    ! neutral%new / %ion_deposit / %update, the level fractions, rho_ion
    ! history, and the ADK coefficients of a level
    function qpad_neutral_new(ctx, cfg, sp_id) bind(C, name="qpad_neutral_new") result(rc)
      import :: c_int, c_int32_t, c_ptr, qpad_neutral_cfg
      type(c_ptr), value :: ctx
      type(qpad_neutral_cfg), intent(in) :: cfg
      integer(c_int32_t), intent(out) :: sp_id
      integer(c_int) :: rc
    end function

    function qpad_neutral_ion_deposit(ctx, sp) bind(C, name="qpad_neutral_ion_deposit") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int) :: rc
    end function

    function qpad_neutral_update(ctx, sp) bind(C, name="qpad_neutral_update") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      integer(c_int) :: rc
    end function

    function qpad_get_neutral_levels(ctx, sp, out) bind(C, name="qpad_get_neutral_levels") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_get_ion_density(ctx, sp, s, out) bind(C, name="qpad_get_ion_density") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp, s
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_neutral_adk(element, level, out) bind(C, name="qpad_neutral_adk") result(rc)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: element, level
      real(c_double), intent(out) :: out(3)
      integer(c_int) :: rc
    end function

    function qpad_get_species_qn(ctx, sp, out) bind(C, name="qpad_get_species_qn") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_set_species_qn(ctx, sp, in) bind(C, name="qpad_set_species_qn") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sp
      real(c_double), intent(in) :: in(*)
      integer(c_int) :: rc
    end function

    function qpad_beam_count(ctx, ib, npp) bind(C, name="qpad_beam_count") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ib
      integer(c_int64_t), intent(out) :: npp
      integer(c_int) :: rc
    end function

    function qpad_get_beam(ctx, ib, x1, x2, x3, p1, p2, p3, q) bind(C, name="qpad_get_beam") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ib
      type(c_ptr), value :: x1, x2, x3, p1, p2, p3, q
      integer(c_int) :: rc
    end function

    function qpad_set_beam(ctx, ib, n, x1, x2, x3, p1, p2, p3, q) bind(C, name="qpad_set_beam") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ib
      integer(c_int64_t), value :: n
      type(c_ptr), value :: x1, x2, x3, p1, p2, p3, q
      integer(c_int) :: rc
    end function

    function qpad_kernel_time(ctx, which, ms, launches) bind(C, name="qpad_kernel_time") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      real(c_double), intent(out) :: ms
      integer(c_int64_t), intent(out) :: launches
      integer(c_int) :: rc
    end function

    function qpad_kernel_time_reset(ctx, enable) bind(C, name="qpad_kernel_time_reset") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
      integer(c_int) :: rc
    end function

    function qpad_iteration_policy(ctx, adaptive, margin) bind(C, name="qpad_iteration_policy") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: adaptive, margin
      integer(c_int) :: rc
    end function

    ! ctxs: c_loc of a type(c_ptr) array of n contexts; batch: the new handle
    function qpad_get_launch_shape(ctx, shape) bind(C, name="qpad_get_launch_shape") result(rc)
      import :: c_int, c_ptr, qpad_launch_shape
      type(c_ptr), value :: ctx
      type(qpad_launch_shape), intent(out) :: shape
      integer(c_int) :: rc
    end function

    function qpad_set_launch_shape(ctx, shape) bind(C, name="qpad_set_launch_shape") result(rc)
      import :: c_int, c_ptr, qpad_launch_shape
      type(c_ptr), value :: ctx
      type(qpad_launch_shape), intent(in) :: shape
      integer(c_int) :: rc
    end function

    ! ms_out: c_loc of real(c_double) :: ms(3), or c_null_ptr
    function qpad_tune_launch(ctx, reps, ms_out) bind(C, name="qpad_tune_launch") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, ms_out
      integer(c_int32_t), value :: reps
      integer(c_int) :: rc
    end function

    function qpad_batch_tune(batch, reps) bind(C, name="qpad_batch_tune") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: reps
      integer(c_int) :: rc
    end function

    function qpad_batch_create(ctxs, n, batch) bind(C, name="qpad_batch_create") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctxs
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: batch
      integer(c_int) :: rc
    end function

    function qpad_batch_destroy(batch) bind(C, name="qpad_batch_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int) :: rc
    end function

    ! mode: 0 eager (host thread per context), 1 graph replay, 2 eager (one host thread)
    function qpad_batch_mode(batch, mode) bind(C, name="qpad_batch_mode") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: mode
      integer(c_int) :: rc
    end function

    function qpad_batch_step3d(batch, total_iters) bind(C, name="qpad_batch_step3d") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), intent(out) :: total_iters(*)
      integer(c_int) :: rc
    end function

    function qpad_seednet_create(ctx, cfg) bind(C, name="qpad_seednet_create") result(rc)
      import :: c_int, c_ptr, qpad_seednet_cfg
      type(c_ptr), value :: ctx
      type(qpad_seednet_cfg), intent(in) :: cfg
      integer(c_int) :: rc
    end function

    function qpad_seednet_get_info(ctx, info) bind(C, name="qpad_seednet_get_info") result(rc)
      import :: c_int, c_ptr, qpad_seednet_info
      type(c_ptr), value :: ctx
      type(qpad_seednet_info), intent(out) :: info
      integer(c_int) :: rc
    end function

    function qpad_seednet_mode(ctx, seed, train) bind(C, name="qpad_seednet_mode") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: seed, train
      integer(c_int) :: rc
    end function

    function qpad_seednet_get_params(ctx, theta, n) bind(C, name="qpad_seednet_get_params") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(out) :: theta(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function

    function qpad_seednet_set_params(ctx, theta, n) bind(C, name="qpad_seednet_set_params") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: theta(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function

    function qpad_seednet_losses(ctx, first, n, out) bind(C, name="qpad_seednet_losses") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first, n
      real(c_float), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function qpad_seednet_seed(ctx, j) bind(C, name="qpad_seednet_seed") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: j
      integer(c_int) :: rc
    end function

    function qpad_seednet_learn(ctx, j) bind(C, name="qpad_seednet_learn") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: j
      integer(c_int) :: rc
    end function

    ! x, t, y, grad: pass c_loc of the arrays, or c_null_ptr to omit
    function qpad_seednet_eval(ctx, x, t, npos, y, loss, grad, update) bind(C, name="qpad_seednet_eval") &
        result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, x, t, y, loss, grad
      integer(c_int32_t), value :: npos, update
      integer(c_int) :: rc
    end function

  end interface

  public :: qpad_last_error, qpad_abi_version, qpad_create, qpad_destroy, qpad_get_grid, &
    qpad_species_new, qpad_species_renew, qpad_species_qdp, qpad_species_qdp_noalloc, &
    qpad_species_charge_history, qpad_species_cbq, qpad_get_species_charge, qpad_species_amjdp, &
    qpad_species_push, qpad_species_push_clamped, qpad_species_set_profile, qpad_beam_lattice_box, &
    qpad_beam_new_std_density, qpad_set_subcycling, qpad_exp_fac_max, qpad_last_subcycles, &
    qpad_species_interp_psi, qpad_step_iterations, qpad_field_solve_psi, qpad_field_solve_ez, &
    qpad_set_laser, qpad_laser_slice, qpad_deposit_chi, &
    qpad_field_solve_et, qpad_field_solve_bz, qpad_field_solve_bt_beam, &
    qpad_field_solve_djdxi, qpad_field_solve_bt_iter, qpad_field_add_b, &
    qpad_convergence_tester, qpad_beam_qdp, qpad_beam_push, qpad_beam_new, qpad_beam_new_std, qpad_beam_new_arrays, &
    qpad_step3d_begin, qpad_slice, qpad_step3d_end, qpad_step3d, &
    qpad_get_field, qpad_set_field, qpad_get_hist, qpad_set_hist, &
    qpad_species_count, qpad_get_species, qpad_set_species, &
    qpad_get_species_qn, qpad_set_species_qn, qpad_beam_count, qpad_get_beam, &
    qpad_neutral_new, qpad_neutral_ion_deposit, qpad_neutral_update, qpad_get_neutral_levels, &
    qpad_get_ion_density, qpad_neutral_adk, &
    qpad_set_beam, qpad_kernel_time, qpad_kernel_time_reset, qpad_iteration_policy, &
    qpad_seednet_create, qpad_seednet_get_info, qpad_seednet_mode, qpad_seednet_get_params, &
    qpad_seednet_set_params, qpad_seednet_losses, qpad_seednet_seed, qpad_seednet_learn, &
    qpad_seednet_eval, qpad_batch_create, qpad_batch_destroy, qpad_batch_mode, qpad_batch_step3d, &
    qpad_get_launch_shape, qpad_set_launch_shape, qpad_tune_launch, qpad_batch_tune

contains

  ! the library's message for the last failed call of this thread
  function qpad_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
      function c_strlen(str) bind(C, name="strlen") result(len)
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
        integer(c_size_t) :: len
      end function
    end interface
    p = qpad_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate (character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function qpad_error_message

  ! QPAD-style error handling: a nonzero code goes to the caller's error
  ! routine (write_err in QPAD) with the library's message
  subroutine qpad_hip_check(rc, where, on_error)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    interface
      subroutine on_error(msg)
        character(len=*), intent(in) :: msg
      end subroutine
    end interface
    if (rc /= QPAD_OK) call on_error('HIP: '//where//': '//qpad_error_message())
  end subroutine qpad_hip_check

end module qpad_hip_mod
