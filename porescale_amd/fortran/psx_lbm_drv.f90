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
! a block appended to ind_red_blue.  settings_19 is the reference's
! (:1442-1514): the run's parameters, node and table counts of the fluid
! and of the solid lattice (its bond/angle table sizes and z-layer groups,
! nthread = 1 as the reference fixes it, :87).
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
!                  1 = FP32 storage with FP64 arithmetic, 2 = the same with
!                  the colour fraction to 33 bits
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
  integer(c_int64_t) :: lcounts(3), layb(3), lib(3), lia(3)
  integer(c_int32_t) :: layers(2)

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
  select case (prec)
  case (1)
     p%precision = PSX_PREC_F32
  case (2)
     p%precision = PSX_PREC_F32X
  case default
     p%precision = PSX_PREC_F64
  end select
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

  call write_settings_19()

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

  ! settings_19 in the reference's layout and list-directed formats
  ! (lsm_lbm_static_alpha_beta_multipore.f:1442-1514): the integers it writes
  ! are default integers (nbo, nna integer(8)) and the reals real(8), as
  ! there.
  subroutine write_settings_19()
    integer :: uo, k
    real(8) :: tr(19), tb(19), cb2, alf, bet
    cb2 = rr0 * cr2 / rb0
    tr(1) = 1.d0 - 2.d0 * cr2 * dt * dt
    tr(2:7) = cr2 * dt * dt * (1.d0 / 6.d0)
    tr(8:19) = cr2 * dt * dt * (1.d0 / 12.d0)
    tb(1) = 1.d0 - 2.d0 * cb2 * dt * dt
    tb(2:7) = cb2 * dt * dt * (1.d0 / 6.d0)
    tb(8:19) = cb2 * dt * dt * (1.d0 / 12.d0)
    alf = dl * Em / (5.d0 - 10.d0 * nu)
    bet = dl * dl * dl * Em * (4 * nu - 1) / (20.d0 * (2 * nu * nu + nu - 1))
    call psx_check(psx_lsm_table_counts(int(ncx, c_int32_t), int(ncy, c_int32_t), int(ncz, c_int32_t), ind, &
         -1_c_int32_t, 1_c_int32_t, lcounts, layers, layb, lib, lia), 'psx_lsm_table_counts')
    open(newunit=uo, file='settings_19', status='replace')
    write(uo, *) 'LBM settings'
    write(uo, *) 'Physical parameters'
    write(uo, *) 'iipore = ', iipore
    write(uo, *) 'npores = ', npores
    write(uo, *) 'surface tension parameter sig = ', sig
    write(uo, *) 'wettability parameter wet = ', wet
    write(uo, *) 'initial red fluid density rr0 = ', rr0
    write(uo, *) 'red fluid cinematic viscosity vir = ', vir
    write(uo, *) 'red fluid sound speed square cr2 = ', cr2
    write(uo, *) 'red fluid lattice weigts tr = ', tr
    write(uo, *) 'red fluid body force bfr = ', bfr
    write(uo, *) 'red fluid collision vector Sr = ', Sr
    write(uo, *) 'initial blue fluid density rb0 = ', rb0
    write(uo, *) 'blue fluid cinematic viscosity vib = ', vib
    write(uo, *) 'blue fluid sound speed square cb2 = ', cb2
    write(uo, *) 'blue fluid lattice weigts tb = ', tb
    write(uo, *) 'blue fluid body force bfb = ', bfb
    write(uo, *) 'blue fluid collision vector Sb = ', Sb
    write(uo, *) 'Computational domain properties'
    write(uo, *) 'Unit cell size xyz = ', ncx, 'x', ncy, 'x', ncz
    write(uo, *) 'number of liquid points nl = ', int(counts(1))
    write(uo, *) 'simple propagation entries nps = ', int(counts(2))
    write(uo, *) 'bounce-back boundary condition entries npb2 = ', int(counts(3))
    write(uo, *) 'bounce-back boundary condition entries npb4 = ', int(counts(4))
    write(uo, *) 'multirefletion boundary condition entries npm = ', 0
    write(uo, *) 'linear size for Reynolds number lsize = ', lsize
    write(uo, *) 'Lattice Boltzmann algorithm parameters'
    write(uo, *) 'interface width parameter B = ', B
    write(uo, *) 'direction of the body force idir = ', idir
    write(uo, *) 'number of steps in the main loop itm = ', itm
    write(uo, *) 'number of steps in the inner loop itw = ', itw
    write(uo, *) 'save parameter its = ', its
    write(uo, *) 'permeability convergence test parameter eps = ', eps
    write(uo, *) 'LBM type (1 - interpolation) mtype = ', mtype
    write(uo, *) 'read initial state data from archiv? icon = ', icon
    write(uo, *) 'small number to supress small distributions ee = ', ee
    write(uo, *) 'LSM settings'
    write(uo, *) 'Physical parameters'
    write(uo, *) 'Physical parameters'
    write(uo, *) 'Young modulus Em = ', Em
    write(uo, *) 'Poissons ratio nu = ', nu
    write(uo, *) 'linear spring parameter alf = ', alf
    write(uo, *) 'angular spring parameter bet = ', bet
    write(uo, *) 'solid medium density rs = ', rs
    write(uo, *) 'integration time step dt = ', dt
    write(uo, *) 'lattice unit step dl = ', dl
    write(uo, *) 'Computational domain properties'
    write(uo, *) 'x unit cell size ncx = ', ncx
    write(uo, *) 'y unit cell size ncy = ', ncy
    write(uo, *) 'z unit cell size ncz = ', ncz
    write(uo, *) 'number of solid points ns = ', int(lcounts(3))
    write(uo, *) 'LSM algorithm parameters'
    write(uo, *) 'nbo = ', lcounts(1)
    write(uo, *) 'nna = ', lcounts(2)
    write(uo, *) 'layers = ', (int(layers(k)), k = 1, 2)
    write(uo, *) 'layb = ', (int(layb(k)), k = 1, 3)
    write(uo, *) 'lib = ', (int(lib(k)), k = 1, 3)
    write(uo, *) 'lia = ', (int(lia(k)), k = 1, 3)
    close(uo)
  end subroutine write_settings_19

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
