! Two-phase LBM driver on the GPU library: the LBM part of
! LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f with its input and output
! files unchanged.
!
! Reads INPUT_PARAMETERS (lines 1-30 as the reference reads them,
! lsm_lbm_static_alpha_beta_multipore.f:435-469), velocities19 and ind; with
! icon = 1 restarts from iteration2 (:2845-2850).  Runs the 2 pre-steps and the
! itm x itw majority-rule steps (:3454-3947), rewrites iteration2 after every
! itw steps (:3951-3961) and appends one output.m row per check_state
! (:3274-3282).  Every its-th iteration it writes the fluid part of
! save_results (:2067-2155): velocity_field, color_field, pressure<it1> and
! a block appended to ind_red_blue.  settings_19 receives the node and table
! counts.
!
! When the cube_bon/cube_ang/cube_sbond/cube_abond files are present (the
! reference's run directory has them), the driver runs the reference's
! coupled model as it does: after the pre-steps every step also moves the
! solid lattice by the fluid's momentum exchange and lsm_step (:3849-3888),
! save_results adds the LSM_* files (:2176-2207) and every check appends
! calc_ab's block to alpha_beta (:3987-3991, 4447-4466).
!
! Optional lines after line 30 (ignored by the reference, which stops reading
! at line 30) select the GPU options:
!   31  precision  0 = FP64 populations (bit-identical to the reference),
!                  1 = FP32 storage with FP64 arithmetic
!   32  collision  0 = the reference's TRT (collmatrTRT), 1 = TRT with the
!                  magic parameter of line 33, 2 = MRT (viscous rates from
!                  vir/vib, free rates 1.19, 1.4, 1.2, 1.4, 1.98)
!   33  lambda     TRT magic parameter for collision = 1
program psx_lbm_drv
  use porescale
  implicit none
  type(psx_lbm_params) :: p
  type(psx_lbm_diag) :: d
  type(c_ptr) :: h
  real(8) :: cr2, rr0, rb0, B, sig, vir, vib, wet, eps, sat, ee, lsize
  real(8) :: bfr(3), bfb(3), Em, nu, rs, dt, dl, lam
  integer :: idir, itm, itw, its, mtype, icon, iipore, npores, ncx, ncy, ncz
  integer :: prec, coll, ios, u, it1, itms, j, i, is
  integer(c_int32_t), allocatable :: ind(:, :, :)
  real(8), allocatable :: fr(:, :), fb(:, :)
  integer(c_int64_t) :: counts(5), nl
  real(8) :: Sr(19), Sb(19)
  type(c_ptr) :: cpl
  logical :: coupled
  integer(c_int32_t), allocatable :: cbon(:), cang(:)
  real(c_double), allocatable :: csb(:), cab(:)

  open(newunit=u, file='INPUT_PARAMETERS', status='old', iostat=ios)
  if (ios /= 0) then
     write(*, *) 'ERROR! cannot open INPUT_PARAMETERS'
     stop 1
  end if
  read(u, *) cr2
  read(u, *) rr0
  read(u, *) rb0
  read(u, *) B
  read(u, *) sig
  read(u, *) vir
  read(u, *) vib
  read(u, *) wet
  read(u, *) idir
  read(u, *) itm
  read(u, *) itw
  read(u, *) its
  read(u, *) eps
  read(u, *) mtype
  read(u, *) icon
  read(u, *) sat
  read(u, *) iipore
  read(u, *) npores
  read(u, *) ee
  read(u, *) lsize
  read(u, *) (bfr(i), i = 1, 3)
  read(u, *) (bfb(i), i = 1, 3)
  read(u, *) ncx
  read(u, *) ncy
  read(u, *) ncz
  read(u, *) Em
  read(u, *) nu
  read(u, *) rs
  read(u, *) dt
  read(u, *) dl
  prec = 0
  coll = 0
  lam = 3.d0 / 16.d0
  read(u, *, iostat=ios) prec
  if (ios /= 0) prec = 0
  if (ios == 0) then
     read(u, *, iostat=ios) coll
     if (ios /= 0) coll = 0
  end if
  if (ios == 0 .and. coll == 1) then
     read(u, *, iostat=ios) lam
     if (ios /= 0) then
        write(*, *) 'ERROR! collision = 1 needs lambda on line 33'
        stop 1
     end if
  end if
  close(u)
  if (mtype /= 0) then
     write(*, *) 'ERROR! only the majority rule (mtype = 0) runs on the GPU'
     stop 1
  end if

  call psx_require_velocities19('velocities19')
  allocate(ind(ncx, ncy, ncz))
  call psx_read_ind('ind', ncx, ncy, ncz, ind)

  call psx_lbm_defaults(p)
  p%cr2 = cr2; p%rr0 = rr0; p%rb0 = rb0; p%B = B; p%sig = sig
  p%vir = vir; p%vib = vib; p%wet = wet; p%idir = idir; p%iipore = iipore
  p%ee = ee; p%lsize = lsize; p%bfr = bfr; p%bfb = bfb; p%dt = dt
  p%precision = merge(PSX_PREC_F32, PSX_PREC_F64, prec == 1)
  select case (coll)
  case (0)
     p%collision = PSX_COLL_REF_TRT
  case (1)
     p%collision = PSX_COLL_TRT
     p%lambda = lam
  case (2)
     p%collision = PSX_COLL_MRT
     call mrt(vir, dt, p%Sr)
     call mrt(vib, dt, p%Sb)
  case default
     write(*, *) 'ERROR! collision must be 0, 1 or 2'
     stop 1
  end select

  call psx_check(psx_lbm_create(p, int(ncx, c_int32_t), int(ncy, c_int32_t), int(ncz, c_int32_t), &
       ind, h), 'psx_lbm_create')
  call psx_check(psx_lbm_counts(h, counts), 'psx_lbm_counts')
  nl = counts(1)
  call psx_check(psx_lbm_rates(h, Sr, Sb), 'psx_lbm_rates')
  allocate(fr(19, nl), fb(19, nl))

  if (icon == 0) then
     call psx_check(psx_lbm_init_random(h, sat), 'psx_lbm_init_random')
     itms = 1
     open(newunit=u, file='output.m', status='replace')
     write(u, *) '% 1 - iteration'
     write(u, *) '% 2 - red  permeability'
     write(u, *) '% 3 - blue permeability'
     write(u, *) '% 4 - red capillary number, volume average'
     write(u, *) '% 5 - blue capillary number, volume average'
     write(u, *) '% 6 - red capillary number, phase average'
     write(u, *) '% 7 - blue capillary number, phase average'
     write(u, *) '% 8 - capillary pressure P_r - P_b, phase average'
     write(u, *) '% 9 - capillary pressure P_red, phase average'
     write(u, *) '% 10 - capillary pressure P_blue, phase average'
     write(u, *) '% 11 - saturation red = number of red/number'
     write(u, *) '% 12 - saturation blue = number of blue/number'
     write(u, *) 'out = ['
     close(u)
  else
     open(newunit=u, file='iteration2', status='old', iostat=ios)
     if (ios /= 0) then
        write(*, *) 'ERROR! icon = 1 needs iteration2'
        stop 1
     end if
     read(u, *) itms
     read(u, *) ((fr(j, i), j = 1, 19), i = 1, nl)
     read(u, *) ((fb(j, i), j = 1, 19), i = 1, nl)
     close(u)
     itms = itms + 1
     call psx_check(psx_lbm_set_state(h, fr, fb), 'psx_lbm_set_state')
  end if

  open(newunit=u, file='settings_19', status='replace')
  write(u, *) 'LBM settings (GPU library)'
  write(u, *) 'iipore = ', iipore
  write(u, *) 'npores = ', npores
  write(u, *) 'surface tension parameter sig = ', sig
  write(u, *) 'wettability parameter wet = ', wet
  write(u, *) 'red fluid collision vector Sr = ', Sr
  write(u, *) 'blue fluid collision vector Sb = ', Sb
  write(u, *) 'number of fluid nodes nl = ', counts(1)
  write(u, *) 'streaming links nps = ', counts(2)
  write(u, *) 'edge bounce-back links npb2 = ', counts(3)
  write(u, *) 'face bounce-back links npb4 = ', counts(4)
  write(u, *) 'precision = ', merge('FP32', 'FP64', prec == 1)
  close(u)

  coupled = read_cube_tables()
  if (coupled) call psx_check(psx_cpl_create(h, Em, nu, dl, rs, cbon, cang, csb, cab, cpl), 'psx_cpl_create')

  ! two pre-steps (lsm_lbm_static_alpha_beta_multipore.f:3454-3621)
  call psx_check(psx_lbm_step(h, 2_c_int64_t), 'psx_lbm_step')
  is = 0
  do it1 = itms, itm
     is = is + 1
     if (coupled) then
        call psx_check(psx_cpl_step(cpl, int(itw, c_int64_t)), 'psx_cpl_step')
     else
        call psx_check(psx_lbm_step(h, int(itw, c_int64_t)), 'psx_lbm_step')
     end if
     call psx_check(psx_lbm_get_state(h, fr, fb), 'psx_lbm_get_state')
     open(newunit=u, file='iteration2', status='replace')
     write(u, *) it1
     write(u, *) ((fr(j, i), j = 1, 19), i = 1, nl)
     write(u, *) ((fb(j, i), j = 1, 19), i = 1, nl)
     close(u)
     if (is == its) then
        is = 0
        call save_fields(it1)
        if (coupled) call save_solid(it1)
     end if
     call psx_check(psx_lbm_check_state(h, int(it1, c_int64_t), int(itw, c_int64_t), d), &
          'psx_lbm_check_state')
     if (d%isneg /= 0) write(*, *) 'ATTENTION!!!, there are negative populations'
     open(newunit=u, file='output.m', position='append')
     write(u, 1984) it1 * itw, d%outm(2:12)
     close(u)
     if (coupled) call alpha_beta()
  end do
1984 format(1X, I9, 11(1X, E15.8), 1X, ';')
  open(newunit=u, file='output.m', position='append')
  write(u, *) '];'
  close(u)
  if (coupled) call psx_cpl_destroy(cpl)
  call psx_lbm_destroy(h)

contains

  logical function read_cube_tables()
    integer :: v, st
    read_cube_tables = .false.
    allocate(cbon(3 * 24), cang(6 * 72), csb(24), cab(72))
    open(newunit=v, file='cube_bon', status='old', iostat=st)
    if (st /= 0) return
    read(v, *, iostat=st) cbon
    close(v)
    if (st /= 0) return
    open(newunit=v, file='cube_ang', status='old', iostat=st)
    if (st /= 0) return
    read(v, *, iostat=st) cang
    close(v)
    if (st /= 0) return
    open(newunit=v, file='cube_sbond', status='old', iostat=st)
    if (st /= 0) return
    read(v, *, iostat=st) csb
    close(v)
    if (st /= 0) return
    open(newunit=v, file='cube_abond', status='old', iostat=st)
    if (st /= 0) return
    read(v, *, iostat=st) cab
    close(v)
    read_cube_tables = st == 0
  end function read_cube_tables

  ! calc_ab's block of alpha_beta (lsm_lbm_static_alpha_beta_multipore.f:
  ! 4447-4466); the tensors are printed divided by the cube count.
  subroutine alpha_beta()
    real(c_double), allocatable :: o(:)
    real(8) :: Cxxxx, Cyyxx, Cxyxy, vcell, rncub
    integer :: uo
    allocate(o(38 + npores))
    call psx_check(psx_cpl_calc_ab(cpl, int(npores, c_int32_t), 0.d0, o), 'psx_cpl_calc_ab')
    Cxxxx = Em * (nu - 1.d0) / (2.d0 * nu * nu + nu - 1.d0)
    Cyyxx = Em * nu / ((1.d0 + nu) * (1.d0 - 2.d0 * nu))
    Cxyxy = Em / (2.d0 * (1.d0 + nu))
    vcell = ncx * ncy * ncz
    rncub = o(38)
    open(newunit=uo, file='alpha_beta', position='append')
    write(uo, *) 'Elastic medium properties'
    write(uo, *) 'Ypung modulus Em = ', Em
    write(uo, *) 'Poisson ratio nu = ', nu
    write(uo, *) 'Stiffness coefficients'
    write(uo, *) 'C_xxxx = ', Cxxxx
    write(uo, *) 'C_yyxx = ', Cyyxx
    write(uo, *) 'C_xyxy = ', Cxyxy
    write(uo, *) 'Simulation parameters'
    write(uo, *) 'Cell volume vcell = ', vcell
    write(uo, *) 'Solid part volume(elementary cubes) rncub = ', rncub
    write(uo, *) 'Simulation results'
    write(uo, *) 'averaged stress tensor'
    write(uo, *) 'asigmap = '
    write(uo, *) o(1:9) / rncub
    write(uo, *) 'asigmam = '
    write(uo, *) o(10:18) / rncub
    write(uo, *) 'averaged strain'
    write(uo, *) 'astrain = '
    write(uo, *) o(19:27) / rncub
    write(uo, *) 'epsi = ', o(28)
    write(uo, *) 'alphai = '
    write(uo, *) o(29:37)
    write(uo, *) 'betai  = '
    write(uo, *) o(39:38 + npores)
    write(uo, *) 'pres = ', (1.d0 - 0.99d0) * rr0 * cr2
    close(uo)
  end subroutine alpha_beta

  ! The solid-lattice files of save_results (:2176-2207).
  subroutine save_solid(it)
    integer, intent(in) :: it
    real(c_double), allocatable :: r(:, :), v(:, :), a(:, :), af(:, :)
    integer(c_int32_t), allocatable :: icors(:, :), mijks(:, :, :)
    integer(c_int64_t) :: cc(4)
    integer :: uu(6), ii, jj, kk, l
    character(len=32) :: dname
    real(8) :: rsi, z3(3)
    call psx_check(psx_cpl_counts(cpl, cc), 'psx_cpl_counts')
    allocate(r(3, max(cc(1), 1_c_int64_t)), v(3, max(cc(1), 1_c_int64_t)), a(3, max(cc(1), 1_c_int64_t)), &
         af(3, max(cc(1), 1_c_int64_t)), icors(3, max(cc(1), 1_c_int64_t)), mijks(ncx, ncy, ncz))
    call psx_check(psx_cpl_get_state(cpl, r, v, a, af), 'psx_cpl_get_state')
    call psx_check(psx_cpl_nodes(cpl, icors, mijks), 'psx_cpl_nodes')
    rsi = 1.d0 / rs
    z3 = 0.d0
    write(dname, '(A, I0)') 'LSM_disf', it
    open(newunit=uu(1), file='LSM_coord_field', status='replace')
    open(newunit=uu(2), file=trim(dname), status='replace')
    open(newunit=uu(3), file='LSM_velocity_field', status='replace')
    open(newunit=uu(4), file='LSM_force_field', status='replace')
    open(newunit=uu(5), file='LSM_icors', status='replace')
    open(newunit=uu(6), file='LSM_fluid_force', status='replace')
    do kk = 1, ncz
       do jj = 1, ncy
          do ii = 1, ncx
             l = mijks(ii, jj, kk)
             if (l /= 0) then
                write(uu(1), *) r(:, l)
                write(uu(2), *) r(:, l)
                write(uu(3), *) v(:, l)
                write(uu(4), *) a(:, l) / rsi
                write(uu(5), *) icors(:, l)
                write(uu(6), *) af(:, l) / rsi
             else
                write(uu(2), *) z3
                write(uu(3), *) z3
                write(uu(4), *) z3
                write(uu(6), *) z3
             end if
          end do
       end do
    end do
    do l = 1, 6
       close(uu(l))
    end do
  end subroutine save_solid

  ! The fluid files of save_results (lsm_lbm_static_alpha_beta_multipore.f:
  ! 2067-2155) in their list-directed layout; the pressure file is named
  ! pressure<it1> and ind_red_blue receives one more block.
  subroutine save_fields(it)
    integer, intent(in) :: it
    real(c_double), allocatable :: uv(:, :), col(:), pre(:)
    integer(c_int32_t), allocatable :: ph(:, :, :)
    integer :: uo, a, nv, jj, kk
    character(len=32) :: fname, rowfmt
    nv = ncx * ncy * ncz
    allocate(uv(3, nv), col(nv), pre(nv), ph(ncx, ncy, ncz))
    call psx_check(psx_lbm_fields(h, uv, col, pre, ph), 'psx_lbm_fields')
    open(newunit=uo, file='velocity_field', status='replace')
    do a = 1, nv
       write(uo, *) uv(:, a)
    end do
    close(uo)
    open(newunit=uo, file='color_field', status='replace')
    do a = 1, nv
       write(uo, *) col(a)
    end do
    close(uo)
    write(fname, '(A, I0)') 'pressure', it
    open(newunit=uo, file=trim(fname), status='replace')
    do a = 1, nv
       write(uo, *) pre(a)
    end do
    close(uo)
    write(rowfmt, '(A, I0, A)') '(', ncx, 'I2)'
    open(newunit=uo, file='ind_red_blue', position='append')
    write(uo, 4579) it * itw
    write(uo, *) 'a=['
    do kk = 1, ncz
       do jj = 1, ncy
          write(uo, rowfmt) ph(:, jj, kk)
       end do
    end do
    write(uo, *) '];'
    write(uo, *) 'clm'
    close(uo)
4579 format(///, 5X, 'ITER=', I7, ';')
  end subroutine save_fields

  subroutine mrt(visc, dt_, S)
    real(8), intent(in) :: visc, dt_
    real(c_double), intent(out) :: S(19)
    real(8) :: sv
    sv = 1.d0 / (3.d0 * visc * dt_ + 0.5d0)
    S = [0.d0, 1.19d0, 1.4d0, 0.d0, 1.2d0, 0.d0, 1.2d0, 0.d0, 1.2d0, sv, 1.4d0, sv, 1.4d0, &
         sv, sv, sv, 1.98d0, 1.98d0, 1.98d0]
  end subroutine mrt
end program psx_lbm_drv
