! ISO_C_BINDING interface to libporescale (include/porescale.h).
!
! The derived types mirror the C structs field by field (bind(C) gives them
! the C layout); every function returns 0 on success or a negative PSX_E*
! code, and psx_message() returns the library's last error text.  Arrays
! are passed in Fortran order: ind(ncx,ncy,ncz), fr/fb(19,nl), u(3,...).
module porescale
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: PSX_OK = 0, PSX_EINVAL = -1, PSX_EHIP = -2, &
       PSX_ENOMEM = -3, PSX_ESTATE = -4, PSX_ECONV = -5
  integer(c_int32_t), parameter :: PSX_COLL_REF_TRT = 0, PSX_COLL_TRT = 1, PSX_COLL_MRT = 2
  integer(c_int32_t), parameter :: PSX_PREC_F64 = 0, PSX_PREC_F32 = 1, PSX_PREC_F32X = 2
  integer(c_int32_t), parameter :: PSX_XX = 0, PSX_YY = 1, PSX_ZZ = 2, PSX_YZ = 3, &
       PSX_ZX = 4, PSX_XY = 5

  type, bind(C) :: psx_lbm_params
     real(c_double) :: cr2, rr0, rb0, B, sig, vir, vib, wet
     integer(c_int32_t) :: idir, iipore
     real(c_double) :: ee, lsize
     real(c_double) :: bfr(3), bfb(3)
     real(c_double) :: dt
     integer(c_int32_t) :: collision
     real(c_double) :: lambda
     real(c_double) :: Sr(19), Sb(19)
     integer(c_int32_t) :: precision
     integer(c_int32_t) :: reserved(7)
  end type psx_lbm_params

  type, bind(C) :: psx_lbm_diag
     real(c_double) :: permr, permb
     real(c_double) :: Qr(3), Qb(3)
     real(c_double) :: um(3), umr(3), umb(3)
     real(c_double) :: tfr(3), tfb(3)
     real(c_double) :: nr, nb
     real(c_double) :: tmr, tmb, tm
     real(c_double) :: Cpr, Cpb
     real(c_double) :: Re
     integer(c_int32_t) :: isneg, negdens_r, negdens_b, pad_
     real(c_double) :: outm(12)
  end type psx_lbm_diag

  type, bind(C) :: psx_lsm_params
     real(c_double) :: alf1, bet1, alf2, bet2
     real(c_double) :: rs, dt, gstrain
     integer(c_int32_t) :: precision
     integer(c_int32_t) :: reserved(7)
  end type psx_lsm_params

  interface
     function psx_last_error() bind(C, name='psx_last_error')
       import :: c_ptr
       type(c_ptr) :: psx_last_error
     end function psx_last_error

     function psx_device_info(name, name_len, num_cu, hbm_gb) bind(C, name='psx_device_info')
       import :: c_int, c_int32_t, c_char, c_double
       integer(c_int) :: psx_device_info
       character(kind=c_char) :: name(*)
       integer(c_int32_t), value :: name_len
       integer(c_int32_t) :: num_cu
       real(c_double) :: hbm_gb
     end function psx_device_info

     ! ------------------------------------------------------------- LBM --
     function psx_lbm_create(p, nx, ny, nz, ind, h) bind(C, name='psx_lbm_create')
       import :: c_int, c_int32_t, c_ptr, psx_lbm_params
       integer(c_int) :: psx_lbm_create
       type(psx_lbm_params), intent(in) :: p
       integer(c_int32_t), value :: nx, ny, nz
       integer(c_int32_t), intent(in) :: ind(*)
       type(c_ptr), intent(out) :: h
     end function psx_lbm_create

     subroutine psx_lbm_destroy(h) bind(C, name='psx_lbm_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine psx_lbm_destroy

     function psx_lbm_counts(h, counts) bind(C, name='psx_lbm_counts')
       import :: c_int, c_int64_t, c_ptr
       integer(c_int) :: psx_lbm_counts
       type(c_ptr), value :: h
       integer(c_int64_t) :: counts(5)
     end function psx_lbm_counts

     function psx_lbm_tables(h, icor, ips, ipb2, ipb4) bind(C, name='psx_lbm_tables')
       import :: c_int, c_int32_t, c_ptr
       integer(c_int) :: psx_lbm_tables
       type(c_ptr), value :: h
       integer(c_int32_t) :: icor(*), ips(*), ipb2(*), ipb4(*)
     end function psx_lbm_tables

     function psx_lbm_rates(h, Sr, Sb) bind(C, name='psx_lbm_rates')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_lbm_rates
       type(c_ptr), value :: h
       real(c_double) :: Sr(19), Sb(19)
     end function psx_lbm_rates

     function psx_lbm_init_random(h, sat) bind(C, name='psx_lbm_init_random')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_lbm_init_random
       type(c_ptr), value :: h
       real(c_double), value :: sat
     end function psx_lbm_init_random

     function psx_lbm_set_state(h, fr, fb) bind(C, name='psx_lbm_set_state')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_lbm_set_state
       type(c_ptr), value :: h
       real(c_double), intent(in) :: fr(*), fb(*)
     end function psx_lbm_set_state

     function psx_lbm_get_state(h, fr, fb) bind(C, name='psx_lbm_get_state')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_lbm_get_state
       type(c_ptr), value :: h
       real(c_double) :: fr(*), fb(*)
     end function psx_lbm_get_state

     function psx_lbm_get_fields(h, rr, rb, cf) bind(C, name='psx_lbm_get_fields')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_lbm_get_fields
       type(c_ptr), value :: h
       real(c_double) :: rr(*), rb(*), cf(*)
     end function psx_lbm_get_fields

     ! save_results fields on the voxel grid (x fastest): u(3, nvox),
     ! color(nvox), pressure(nvox), phase(nvox)
     function psx_lbm_fields(h, u, color, pressure, phase) bind(C, name='psx_lbm_fields')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lbm_fields
       type(c_ptr), value :: h
       real(c_double) :: u(*), color(*), pressure(*)
       integer(c_int32_t) :: phase(*)
     end function psx_lbm_fields

     function psx_lbm_step(h, nsteps) bind(C, name='psx_lbm_step')
       import :: c_int, c_int64_t, c_ptr
       integer(c_int) :: psx_lbm_step
       type(c_ptr), value :: h
       integer(c_int64_t), value :: nsteps
     end function psx_lbm_step

     function psx_lbm_check_state(h, it1, itw, d) bind(C, name='psx_lbm_check_state')
       import :: c_int, c_int64_t, c_ptr, psx_lbm_diag
       integer(c_int) :: psx_lbm_check_state
       type(c_ptr), value :: h
       integer(c_int64_t), value :: it1, itw
       type(psx_lbm_diag) :: d
     end function psx_lbm_check_state

     function psx_lbm_sync(h) bind(C, name='psx_lbm_sync')
       import :: c_int, c_ptr
       integer(c_int) :: psx_lbm_sync
       type(c_ptr), value :: h
     end function psx_lbm_sync

     ! ------------------------------------------------------------- LSM --
     function psx_lsm_create(p, nx, ny, nz, ind, stiff, cube_bon, cube_ang, cube_sbond, &
          cube_abond, h) bind(C, name='psx_lsm_create')
       import :: c_int, c_int32_t, c_ptr, psx_lsm_params
       integer(c_int) :: psx_lsm_create
       type(psx_lsm_params), intent(in) :: p
       integer(c_int32_t), value :: nx, ny, nz
       integer(c_int32_t), intent(in) :: ind(*)
       type(c_ptr), value :: stiff, cube_bon, cube_ang, cube_sbond, cube_abond
       type(c_ptr), intent(out) :: h
     end function psx_lsm_create

     subroutine psx_lsm_destroy(h) bind(C, name='psx_lsm_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine psx_lsm_destroy

     function psx_lsm_counts(h, counts) bind(C, name='psx_lsm_counts')
       import :: c_int, c_int64_t, c_ptr
       integer(c_int) :: psx_lsm_counts
       type(c_ptr), value :: h
       integer(c_int64_t) :: counts(4)
     end function psx_lsm_counts

     ! new spring constants for the same geometry (calibration loops)
     function psx_lsm_set_springs(h, alf1, bet1, alf2, bet2) bind(C, name='psx_lsm_set_springs')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_lsm_set_springs
       type(c_ptr), value :: h
       real(c_double), value :: alf1, bet1, alf2, bet2
     end function psx_lsm_set_springs

     ! host only: bond/angle table sizes and the z-layer partition (lsm_settings)
     function psx_lsm_table_counts(nx, ny, nz, ind, loadcase, nthread, counts, layers, layb, lib, lia) &
          bind(C, name='psx_lsm_table_counts')
       import :: c_int, c_int32_t, c_int64_t
       integer(c_int) :: psx_lsm_table_counts
       integer(c_int32_t), value :: nx, ny, nz, loadcase, nthread
       integer(c_int32_t) :: ind(*)
       integer(c_int64_t) :: counts(3)
       integer(c_int32_t) :: layers(*)
       integer(c_int64_t) :: layb(*), lib(*), lia(*)
     end function psx_lsm_table_counts

     function psx_lsm_get_u(h, loadcase, u) bind(C, name='psx_lsm_get_u')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_get_u
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       real(c_double) :: u(*)
     end function psx_lsm_get_u

     function psx_lsm_verlet(h, loadcase, nsteps, restart) bind(C, name='psx_lsm_verlet')
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int) :: psx_lsm_verlet
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       integer(c_int64_t), value :: nsteps
       integer(c_int32_t), value :: restart
     end function psx_lsm_verlet

     function psx_lsm_reference_nodes(h, loadcase, count, grid, twin, icors) bind(C, name='psx_lsm_reference_nodes')
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int) :: psx_lsm_reference_nodes
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       integer(c_int64_t) :: count
       type(c_ptr), value :: grid, twin, icors
     end function psx_lsm_reference_nodes

     ! iteration2 state (rdat, vdat, adat as (3,ns)) of the Verlet loop
     function psx_lsm_get_verlet(h, loadcase, r, v, a) bind(C, name='psx_lsm_get_verlet')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_get_verlet
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       real(c_double) :: r(*), v(*), a(*)
     end function psx_lsm_get_verlet

     function psx_lsm_set_verlet(h, loadcase, r, v, a) bind(C, name='psx_lsm_set_verlet')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_set_verlet
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       real(c_double) :: r(*), v(*), a(*)
     end function psx_lsm_set_verlet

     function psx_lsm_solve(h, loadcase, rtol, maxit, iters, relres) bind(C, name='psx_lsm_solve')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_solve
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       real(c_double), value :: rtol
       integer(c_int32_t), value :: maxit
       integer(c_int32_t) :: iters
       real(c_double) :: relres
     end function psx_lsm_solve

     function psx_lsm_stress(h, loadcase, asig, moduli) bind(C, name='psx_lsm_stress')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_stress
       type(c_ptr), value :: h
       integer(c_int32_t), value :: loadcase
       real(c_double) :: asig(3, 3), moduli(3, 3)
     end function psx_lsm_stress

     function psx_lsm_moduli(h, rtol, maxit, C, K, G) bind(C, name='psx_lsm_moduli')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_moduli
       type(c_ptr), value :: h
       real(c_double), value :: rtol
       integer(c_int32_t), value :: maxit
       real(c_double) :: C(6, 6), K, G
     end function psx_lsm_moduli

     function psx_lsm_adjoint(h, wK, wG, rtol, maxit, dab, dstiff) bind(C, name='psx_lsm_adjoint')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_lsm_adjoint
       type(c_ptr), value :: h
       real(c_double), value :: wK, wG, rtol
       integer(c_int32_t), value :: maxit
       real(c_double) :: dab(4)
       type(c_ptr), value :: dstiff
     end function psx_lsm_adjoint

     ! coupled LBM -> LSM model (lsm_lbm_static_alpha_beta_multipore.f)
     function psx_cpl_create(lbm, Em, nu, dl, rs, cube_bon, cube_ang, cube_sbond, cube_abond, h) &
          bind(C, name='psx_cpl_create')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_cpl_create
       type(c_ptr), value :: lbm
       real(c_double), value :: Em, nu, dl, rs
       integer(c_int32_t) :: cube_bon(*), cube_ang(*)
       real(c_double) :: cube_sbond(*), cube_abond(*)
       type(c_ptr), intent(out) :: h
     end function psx_cpl_create

     subroutine psx_cpl_destroy(h) bind(C, name='psx_cpl_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine psx_cpl_destroy

     function psx_cpl_counts(h, counts) bind(C, name='psx_cpl_counts')
       import :: c_int, c_int64_t, c_ptr
       integer(c_int) :: psx_cpl_counts
       type(c_ptr), value :: h
       integer(c_int64_t) :: counts(4)
     end function psx_cpl_counts

     function psx_cpl_nodes(h, icors, mijks) bind(C, name='psx_cpl_nodes')
       import :: c_int, c_int32_t, c_ptr
       integer(c_int) :: psx_cpl_nodes
       type(c_ptr), value :: h
       integer(c_int32_t) :: icors(*), mijks(*)
     end function psx_cpl_nodes

     function psx_cpl_step(h, nsteps) bind(C, name='psx_cpl_step')
       import :: c_int, c_int64_t, c_ptr
       integer(c_int) :: psx_cpl_step
       type(c_ptr), value :: h
       integer(c_int64_t), value :: nsteps
     end function psx_cpl_step

     function psx_cpl_get_state(h, r, v, a, af) bind(C, name='psx_cpl_get_state')
       import :: c_int, c_double, c_ptr
       integer(c_int) :: psx_cpl_get_state
       type(c_ptr), value :: h
       real(c_double) :: r(*), v(*), a(*), af(*)
     end function psx_cpl_get_state

     function psx_cpl_calc_ab(h, npores, pres, out) bind(C, name='psx_cpl_calc_ab')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int) :: psx_cpl_calc_ab
       type(c_ptr), value :: h
       integer(c_int32_t), value :: npores
       real(c_double), value :: pres
       real(c_double) :: out(*)
     end function psx_cpl_calc_ab
  end interface

contains

  ! The library's last error message as a Fortran string.
  function psx_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = psx_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function psx_message

  ! Stops the driver with the library's message when rc is an error (the
  ! drivers keep the reference's write-and-stop behaviour; the library never
  ! stops the process itself).
  subroutine psx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= PSX_OK) then
       write(*, *) 'ERROR! ', what, ' failed (', rc, '): ', psx_message()
       stop 1
    end if
  end subroutine psx_check

  ! Default LBM parameters: the reference's TRT collision, FP64 storage.
  subroutine psx_lbm_defaults(p)
    type(psx_lbm_params), intent(out) :: p
    p%cr2 = 1.d0 / 3.d0; p%rr0 = 1.d0; p%rb0 = 1.d0; p%B = 0.7d0; p%sig = 0.01d0
    p%vir = 1.d0 / 6.d0; p%vib = 1.d0 / 6.d0; p%wet = 0.d0
    p%idir = 3; p%iipore = 1; p%ee = 0.d0; p%lsize = 1.d0
    p%bfr = 0.d0; p%bfb = 0.d0; p%dt = 1.d0
    p%collision = PSX_COLL_REF_TRT; p%lambda = 3.d0 / 16.d0
    p%Sr = 0.d0; p%Sb = 0.d0; p%precision = PSX_PREC_F64; p%reserved = 0
  end subroutine psx_lbm_defaults

  ! Reads the reference's velocities19 file and stops unless it is the
  ! D3Q19 set in the order the library uses.
  subroutine psx_require_velocities19(path)
    character(len=*), intent(in) :: path
    integer :: v(3, 19), ios, u
    integer, parameter :: ref(3, 19) = reshape([ &
         0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, &
         1, 1, 0, -1, 1, 0, 1, -1, 0, -1, -1, 0, 1, 0, 1, -1, 0, 1, 1, 0, -1, -1, 0, -1, &
         0, 1, 1, 0, -1, 1, 0, 1, -1, 0, -1, -1], [3, 19])
    open(newunit=u, file=path, status='old', iostat=ios)
    if (ios /= 0) return   ! the set is built into the library
    read(u, '(19i3)', iostat=ios) v(1, :)
    if (ios == 0) read(u, '(19i3)', iostat=ios) v(2, :)
    if (ios == 0) read(u, '(19i3)', iostat=ios) v(3, :)
    close(u)
    if (ios /= 0) then
       write(*, *) 'ERROR! cannot read ', path
       stop 1
    end if
    if (any(v /= ref)) then
       write(*, *) 'ERROR! ', path, ' is not the D3Q19 velocity set of this library'
       stop 1
    end if
  end subroutine psx_require_velocities19

  ! Reads the reference's 'ind' file: per (j,k), k outer, one record of a
  ! blank and ncx single digits.
  subroutine psx_read_ind(path, ncx, ncy, ncz, ind)
    character(len=*), intent(in) :: path
    integer, intent(in) :: ncx, ncy, ncz
    integer(c_int32_t), intent(out) :: ind(ncx, ncy, ncz)
    character(len=64) :: fmt
    integer :: i, j, k, u, ios
    write(fmt, '(a,i0,a)') '(1x,', ncx, 'i1)'
    open(newunit=u, file=path, status='old', iostat=ios)
    if (ios /= 0) then
       write(*, *) 'ERROR! cannot open ', path
       stop 1
    end if
    do k = 1, ncz
       do j = 1, ncy
          read(u, fmt, iostat=ios) (ind(i, j, k), i = 1, ncx)
          if (ios /= 0) then
             write(*, *) 'ERROR! short ind file at record ', j + (k - 1) * ncy
             stop 1
          end if
       end do
    end do
    close(u)
  end subroutine psx_read_ind
end module porescale
