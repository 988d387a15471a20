! qpad_slice_driver -- the QPAD side of the drop-in, as a program: one 3D
! step driven through qpad_hip_mod one procedure at a time, in the order of
! QPAD's simulation%run (simulation_class.f03:294-469): per slice the beam
! B_perp, the charge deposit and psi, Bz from the J_perp predictor, then the
! predictor-corrector iteration (record, add_f1, E, current deposit, dJ/dxi,
! B_perp, Bz, compare), then species2d%cbq, E_spe, E and the J_perp
! predictor, and the plasma push.  QPAD's own array arithmetic (zeroing,
! cbq, the predictor update) is done on host copies of the slice fields, as
! QPAD's Fortran does on its ufield arrays.
!
! usage: qpad_slice_driver <parameters> <output>
!   parameters (list-directed): the qpad_grid components; the species count
!   and each qpad_species_cfg; one qpad_beam_cfg (random Gaussian beam)
!   output (unformatted stream): the grid as read, the iterations of every
!   slice, then the slice fields psi, e, b, b_spe, cu, q_spe after the last
!   slice ([ncomp][nm][nr+2] each) and species 0's particle count and
!   x1, x2, p1, p2, p3, id.
program qpad_slice_driver
  use qpad_hip_mod
  use, intrinsic :: iso_c_binding
  implicit none
  type(qpad_grid) :: g
  type(qpad_species_cfg) :: sc(4)
  type(qpad_beam_cfg) :: bc
  type(c_ptr) :: ctx
  integer(c_int32_t) :: nsp, k, j, l, sid, bid
  integer(c_int64_t) :: npp
  integer :: u, nm, ld, f
  integer(c_int32_t), allocatable :: iters(:)
  integer(c_int32_t), allocatable, target :: pid(:)
  real(c_double), allocatable :: cu(:,:,:), q(:,:,:), dcu(:,:,:), zero3(:,:,:), zero2(:,:,:), out(:,:,:)
  real(c_double), allocatable, target :: x1(:), x2(:), p1(:), p2(:), p3(:), pq(:), pg(:), pp(:)
  real(c_double) :: rel, ab
  character(len=1024) :: inp, outp
  integer(c_int32_t), parameter :: dump(6) = [QPAD_F_PSI, QPAD_F_E, QPAD_F_B, QPAD_F_B_SPE, QPAD_F_CU, &
                                              QPAD_F_Q_SPE]
  integer, parameter :: ncomp(0:11) = [1, 3, 3, 3, 3, 3, 3, 2, 3, 2, 1, 1]

  call get_command_argument(1, inp)
  call get_command_argument(2, outp)
  open(newunit=u, file=trim(inp), status='old', action='read')
  read(u, *) g%nr, g%nxi, g%max_mode, g%dr, g%dxi, g%relax_fac, g%bnd, g%iter_max, g%iter_reltol, &
             g%iter_abstol, g%dt
  read(u, *) nsp
  do k = 1, nsp
    read(u, *) sc(k)%ppc1, sc(k)%ppc2, sc(k)%num_theta, sc(k)%qm, sc(k)%mass, sc(k)%density, sc(k)%uth, &
               sc(k)%neutralized, sc(k)%den_min, sc(k)%seed, sc(k)%push_type
  end do
  read(u, *) bc%total_num, bc%total_charge, bc%qm, bc%mass, bc%gamma, bc%center, bc%sigma, bc%range_lo, &
             bc%range_hi, bc%uth, bc%quiet, bc%evol, bc%seed
  close(u)

  ! simulation%new (simulation_class.f03:77-199)
  call chk(qpad_create(g, ctx), 'qpad_create')
  do k = 1, nsp
    call chk(qpad_species_new(ctx, sc(k), sid), 'qpad_species_new')
  end do
  call chk(qpad_beam_new(ctx, bc, bid), 'qpad_beam_new')

  nm = 2 * g%max_mode + 1
  ld = g%nr + 2
  allocate(iters(g%nxi), cu(0:g%nr+1, nm, 3), q(0:g%nr+1, nm, 1), dcu(0:g%nr+1, nm, 2))
  allocate(zero3(0:g%nr+1, nm, 3), zero2(0:g%nr+1, nm, 2))
  zero3 = 0.0d0
  zero2 = 0.0d0

  ! 3D-step head (simulation_class.f03:294-340)
  call chk(qpad_step3d_begin(ctx), 'qpad_step3d_begin')
  do j = 1, g%nxi
    ! beam B_perp of this slice, plasma charge and psi, Bz from the predictor
    ! (simulation_class.f03:345-368)
    call chk(qpad_field_solve_bt_beam(ctx, j), 'field_b%solve(q_beam)')
    call chk(qpad_set_field(ctx, QPAD_F_Q_SPE, zero3), 'q_spe = 0')
    do k = 1, nsp
      call chk(qpad_species_qdp(ctx, k - 1), 'species2d%qdp')
    end do
    call chk(qpad_field_solve_psi(ctx), 'field_psi%solve')
    do k = 1, nsp
      call chk(qpad_species_interp_psi(ctx, k - 1), 'species2d%interp_psi')
    end do
    call chk(qpad_field_solve_bz(ctx, QPAD_F_B_SPE, QPAD_F_CU), 'field_b%solve(cu)')
    ! predictor-corrector iteration (simulation_class.f03:370-399)
    do l = 1, g%iter_max
      iters(j) = l
      call chk(qpad_convergence_tester(ctx, 0, rel, ab), 'convergence_tester record')
      call chk(qpad_field_add_b(ctx), 'add_f1')
      call chk(qpad_field_solve_ez(ctx, QPAD_F_E, QPAD_F_CU), 'field_e%solve(cu)')
      call chk(qpad_field_solve_et(ctx, QPAD_F_E, QPAD_F_B), 'field_e%solve(b, psi)')
      call chk(qpad_set_field(ctx, QPAD_F_CU, zero3), 'cu = 0')
      call chk(qpad_set_field(ctx, QPAD_F_ACU, zero2), 'acu = 0')
      call chk(qpad_set_field(ctx, QPAD_F_AMU, zero3), 'amu = 0')
      do k = 1, nsp
        call chk(qpad_species_amjdp(ctx, k - 1, g%dxi), 'species2d%amjdp')
      end do
      call chk(qpad_field_solve_djdxi(ctx), 'field_djdxi%solve')
      call chk(qpad_field_solve_bt_iter(ctx), 'field_b%solve(dcu, cu)')
      call chk(qpad_field_solve_bz(ctx, QPAD_F_B_SPE, QPAD_F_CU), 'field_b%solve(cu)')
      call chk(qpad_convergence_tester(ctx, 1, rel, ab), 'convergence_tester compare')
      if (rel < g%iter_reltol .or. ab < g%iter_abstol) exit
    end do
    ! species2d%cbq, E_spe, E, the J_perp predictor and the push
    ! (simulation_class.f03:401-456)
    call chk(qpad_get_field(ctx, QPAD_F_CU, cu), 'get cu')
    call chk(qpad_get_field(ctx, QPAD_F_Q_SPE, q), 'get q_spe')
    q(:, :, 1) = q(:, :, 1) + cu(:, :, 3)
    call chk(qpad_set_field(ctx, QPAD_F_Q_SPE, q), 'set q_spe')
    call chk(qpad_field_add_b(ctx), 'add_f1')
    call chk(qpad_field_solve_et(ctx, QPAD_F_E_SPE, QPAD_F_B_SPE), 'field_e%solve(b_spe, psi)')
    call chk(qpad_field_solve_ez(ctx, QPAD_F_E, QPAD_F_CU), 'field_e%solve(cu)')
    call chk(qpad_field_solve_et(ctx, QPAD_F_E, QPAD_F_B), 'field_e%solve(b, psi)')
    call chk(qpad_get_field(ctx, QPAD_F_DCU, dcu), 'get dcu')
    dcu = g%dxi * dcu
    call chk(qpad_set_field(ctx, QPAD_F_DCU, dcu), 'set dcu')
    cu(:, :, 1:2) = cu(:, :, 1:2) + dcu
    call chk(qpad_set_field(ctx, QPAD_F_CU, cu), 'set cu')
    do k = 1, nsp
      call chk(qpad_species_push(ctx, k - 1, g%dxi), 'species2d%push')
    end do
  end do

  open(newunit=u, file=trim(outp), access='stream', form='unformatted', status='replace', action='write')
  write(u) g%nr, g%nxi, g%max_mode, g%dr, g%dxi, g%relax_fac, g%bnd, g%iter_max, g%iter_reltol, &
           g%iter_abstol, g%dt
  write(u) iters
  do f = 1, size(dump)
    allocate(out(0:g%nr+1, nm, ncomp(dump(f))))
    call chk(qpad_get_field(ctx, dump(f), out), 'get field')
    write(u) out
    deallocate(out)
  end do
  call chk(qpad_species_count(ctx, 0, npp), 'qpad_species_count')
  allocate(x1(npp), x2(npp), p1(npp), p2(npp), p3(npp), pq(npp), pg(npp), pp(npp), pid(npp))
  call chk(qpad_get_species(ctx, 0, c_loc(x1), c_loc(x2), c_loc(p1), c_loc(p2), c_loc(p3), c_loc(pq), &
                            c_loc(pg), c_loc(pp), c_loc(pid)), 'qpad_get_species')
  write(u) npp
  write(u) x1, x2, p1, p2, p3, pid
  close(u)
  call chk(qpad_destroy(ctx), 'qpad_destroy')

contains

  subroutine chk(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    call qpad_hip_check(rc, where, stop_on_error)
  end subroutine chk

  ! QPAD's write_err (sysutil_module.f03:165-183) logs and stops; so does this
  subroutine stop_on_error(msg)
    character(len=*), intent(in) :: msg
    write(0, '(a)') trim(msg)
    error stop 1
  end subroutine stop_on_error

end program qpad_slice_driver
