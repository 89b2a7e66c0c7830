! Lattice spring model driver on the GPU library: the programs
! LSM2S/Cxxxx/XX_P.f (load case xx) and LSM2S/Cyzyz/YZ_P.f (load case yz),
! plus the four other unit-strain cases, with their input files unchanged.
!
! Reads INPUT_PARAMETERS (XX_P.f:364-382: ncx, ncy, ncz, Em1, nu1, Em2, nu2,
! rs, dt, dl, itm, its, gstrain, nthread, icon), velocities19, ind and, when
! present, the cube_bon/cube_ang/cube_sbond/cube_abond element tables
! (XX_P.f:1700-1714).  Spring constants follow XX_P.f:411-414.
!
!   psx_lsm_drv [case] [mode]
!     case  xx (default), yy, zz, yz, zx, xy, or all
!     mode  solve  (default) preconditioned CG to the relative residual of
!                  line 17 (default 1e-10), then one stress_tensor block
!           verlet the reference's damped Verlet relaxation: itm blocks of
!                  its steps, a stress_tensor block, the iteration2 state and
!                  the field files LSM_coord_field, LSM_displ_field,
!                  LSM_velocity_field, LSM_force_field and 'write' after each
!                  (XX_P.f:1377-1609); icon = 1 restarts from iteration2
!                  (XX_P.f:1288-1294)
!   case all solves the six cases and writes the 6x6 stiffness and the Voigt
!   bulk and shear moduli to the file 'moduli'.
!
! Optional line 16 selects the precision (0 = FP64, 1 = FP32 matvec with FP64
! residual refinement) and line 17 the relative residual tolerance.
program psx_lsm_drv
  use porescale
  implicit none
  type(psx_lsm_params) :: p
  type(c_ptr) :: h
  integer :: ncx, ncy, ncz, itm, its, nthread, icon, prec, ios, u, i, iti, lc, nargs
  real(8) :: Em1, nu1, Em2, nu2, rs, dt, dl, gstrain, rtol
  real(8) :: alf1, bet1, alf2, bet2, K, G
  real(8) :: asig(3, 3), moduli(3, 3), C(6, 6)
  integer(c_int32_t) :: iters
  real(8) :: relres
  integer(c_int64_t) :: counts(4), nref
  real(c_double), allocatable :: rr(:, :), vv(:, :), aa(:, :)
  integer(c_int32_t), allocatable, target :: icors(:, :)
  integer(c_int32_t), allocatable :: mijks(:, :, :)
  integer :: j, gdim(3)
  integer, parameter :: pad_axis(0:5) = [1, 2, 3, 2, 3, 1]
  integer(c_int32_t), allocatable :: ind(:, :, :)
  integer(c_int32_t), allocatable, target :: cbon(:), cang(:)
  real(c_double), allocatable, target :: csb(:), cab(:)
  type(c_ptr) :: pbon, pang, psb, pab
  character(len=16) :: arg_case, arg_mode
  character(len=2), parameter :: names(0:5) = ['xx', 'yy', 'zz', 'yz', 'zx', 'xy']

  arg_case = 'xx'
  arg_mode = 'solve'
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, arg_case)
  if (nargs >= 2) call get_command_argument(2, arg_mode)

  open(newunit=u, file='INPUT_PARAMETERS', status='old', iostat=ios)
  if (ios /= 0) then
     write(*, *) 'ERROR! cannot open INPUT_PARAMETERS'
     stop 1
  end if
  read(u, *) ncx
  read(u, *) ncy
  read(u, *) ncz
  read(u, *) Em1
  read(u, *) nu1
  read(u, *) Em2
  read(u, *) nu2
  read(u, *) rs
  read(u, *) dt
  read(u, *) dl
  read(u, *) itm
  read(u, *) its
  read(u, *) gstrain
  read(u, *) nthread
  read(u, *) icon
  prec = 0
  rtol = 1.d-10
  read(u, *, iostat=ios) prec
  if (ios /= 0) prec = 0
  if (ios == 0) then
     read(u, *, iostat=ios) rtol
     if (ios /= 0) rtol = 1.d-10
  end if
  close(u)

  call psx_require_velocities19('velocities19')
  allocate(ind(ncx, ncy, ncz))
  call psx_read_ind('ind', ncx, ncy, ncz, ind)

  alf1 = dl * Em1 / (5.d0 - 10.d0 * nu1)
  bet1 = dl * dl * dl * Em1 * (4 * nu1 - 1) / (20.d0 * (2 * nu1 * nu1 + nu1 - 1))
  alf2 = dl * Em2 / (5.d0 - 10.d0 * nu2)
  bet2 = dl * dl * dl * Em2 * (4 * nu2 - 1) / (20.d0 * (2 * nu2 * nu2 + nu2 - 1))
  p%alf1 = alf1; p%bet1 = bet1; p%alf2 = alf2; p%bet2 = bet2
  p%rs = rs; p%dt = dt; p%gstrain = gstrain
  p%precision = merge(PSX_PREC_F32, PSX_PREC_F64, prec == 1)
  p%reserved = 0

  ! element tables, when the files are there (the library checks them)
  pbon = c_null_ptr; pang = c_null_ptr; psb = c_null_ptr; pab = c_null_ptr
  if (read_cube_tables()) then
     pbon = c_loc(cbon); pang = c_loc(cang); psb = c_loc(csb); pab = c_loc(cab)
  end if

  call psx_check(psx_lsm_create(p, int(ncx, c_int32_t), int(ncy, c_int32_t), int(ncz, c_int32_t), &
       ind, c_null_ptr, pbon, pang, psb, pab, h), 'psx_lsm_create')
  call psx_check(psx_lsm_counts(h, counts), 'psx_lsm_counts')

  if (trim(arg_case) == 'all') then
     call psx_check(psx_lsm_moduli(h, rtol, 1000000_c_int32_t, C, K, G), 'psx_lsm_moduli')
     open(newunit=u, file='moduli', status='replace')
     write(u, *) 'Voigt stiffness C (rows xx yy zz yz zx xy stress, columns load cases)'
     do i = 1, 6
        write(u, '(6(1X,E23.15))') C(i, :)
     end do
     write(u, *) 'bulk modulus K = ', K
     write(u, *) 'shear modulus G = ', G
     close(u)
     call psx_lsm_destroy(h)
     stop
  end if

  lc = -1
  do i = 0, 5
     if (trim(arg_case) == names(i)) lc = i
  end do
  if (lc < 0) then
     write(*, *) 'ERROR! unknown load case ', trim(arg_case)
     stop 1
  end if

  select case (trim(arg_mode))
  case ('solve')
     call psx_check(psx_lsm_solve(h, int(lc, c_int32_t), rtol, 1000000_c_int32_t, iters, relres), &
          'psx_lsm_solve')
     write(*, *) 'CG iterations = ', iters, ' relative residual = ', relres
     call psx_check(psx_lsm_stress(h, int(lc, c_int32_t), asig, moduli), 'psx_lsm_stress')
     call write_block(0, 0)
  case ('verlet')
     call psx_check(psx_lsm_reference_nodes(h, int(lc, c_int32_t), nref, c_null_ptr, c_null_ptr, c_null_ptr), &
          'psx_lsm_reference_nodes')
     allocate(rr(3, max(nref, 1_c_int64_t)), vv(3, max(nref, 1_c_int64_t)), aa(3, max(nref, 1_c_int64_t)))
     allocate(icors(3, max(nref, 1_c_int64_t)))
     call psx_check(psx_lsm_reference_nodes(h, int(lc, c_int32_t), nref, c_null_ptr, c_null_ptr, c_loc(icors)), &
          'psx_lsm_reference_nodes')
     gdim = [ncx, ncy, ncz]
     gdim(pad_axis(lc)) = gdim(pad_axis(lc)) + 2   ! the padded grid (XX_P.f:521)
     allocate(mijks(gdim(1), gdim(2), gdim(3)))
     mijks = 0
     do i = 1, int(nref)
        mijks(icors(1, i), icors(2, i), icors(3, i)) = i
     end do
     call write_settings()
     if (icon == 1) then
        open(newunit=u, file='iteration2', status='old', iostat=ios)
        if (ios /= 0) then
           write(*, *) 'ERROR! icon = 1 needs iteration2'
           stop 1
        end if
        read(u, *) ((rr(j, i), j = 1, 3), i = 1, nref)
        read(u, *) ((vv(j, i), j = 1, 3), i = 1, nref)
        read(u, *) ((aa(j, i), j = 1, 3), i = 1, nref)
        close(u)
        call psx_check(psx_lsm_set_verlet(h, int(lc, c_int32_t), rr, vv, aa), 'psx_lsm_set_verlet')
     end if
     do iti = 1, itm
        call psx_check(psx_lsm_verlet(h, int(lc, c_int32_t), int(its, c_int64_t), &
             merge(1_c_int32_t, 0_c_int32_t, iti == 1 .and. icon /= 1)), 'psx_lsm_verlet')
        call psx_check(psx_lsm_stress(h, int(lc, c_int32_t), asig, moduli), 'psx_lsm_stress')
        call write_block(iti, its)
        call psx_check(psx_lsm_get_verlet(h, int(lc, c_int32_t), rr, vv, aa), 'psx_lsm_get_verlet')
        call write_fields(iti, its + 1)
        open(newunit=u, file='iteration2', status='replace')
        write(u, *) ((rr(j, i), j = 1, 3), i = 1, nref)
        write(u, *) ((vv(j, i), j = 1, 3), i = 1, nref)
        write(u, *) ((aa(j, i), j = 1, 3), i = 1, nref)
        close(u)
     end do
  case default
     write(*, *) 'ERROR! mode must be solve or verlet'
     stop 1
  end select
  call psx_lsm_destroy(h)

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

  ! One block of the reference's stress_tensor file (XX_P.f:1938-1966,
  ! YZ_P.f:1938-1963): the summed cube stress per solid cube and the
  ! effective stiffness coefficients of the load case.
  ! lsm_settings (XX_P.f:1314-1348, YZ_P.f:1309-1341): parameters, sizes on
  ! the padded grid, and the sizes of the reference's bond and angle tables
  ! with their z-layer partition, counted on the host by the library (the
  ! element operator itself uses no tables)
  subroutine write_settings()
    integer :: v, l
    integer(c_int64_t) :: tc(3)
    integer(c_int32_t), allocatable :: lay4(:)
    integer, allocatable :: layers(:)
    integer(8), allocatable :: layb(:), lib(:), lia(:)
    integer(8) :: nbo, nna
    allocate(lay4(2 * nthread), layers(2 * nthread))
    allocate(layb(2 * nthread + 1), lib(2 * nthread + 1), lia(2 * nthread + 1))
    call psx_check(psx_lsm_table_counts(int(ncx, c_int32_t), int(ncy, c_int32_t), int(ncz, c_int32_t), ind, &
         int(lc, c_int32_t), int(nthread, c_int32_t), tc, lay4, layb, lib, lia), 'psx_lsm_table_counts')
    nbo = tc(1); nna = tc(2)
    layers = lay4
    open(newunit=v, file='lsm_settings', status='replace')
    write(v, *) 'Initial program settings'
    write(v, *) 'Physical parameters'
    if (lc >= 3) then  ! the shear program's labels (YZ_P.f:1313-1320)
       write(v, *) 'Young modulus Em1 = ', Em1
       write(v, *) 'Poissons ratio nu1 = ', nu1
       write(v, *) 'linear spring parameter alf1 = ', alf1
       write(v, *) 'angular spring parameter bet1 = ', bet1
       write(v, *) 'Young modulus Em2 = ', Em2
       write(v, *) 'Poissons ratio nu2 = ', nu2
       write(v, *) 'linear spring parameter alf2 = ', alf2
       write(v, *) 'angular spring parameter bet2 = ', bet2
    else
       write(v, *) 'Young modulus Em1 of component solid 1 = ', Em1
       write(v, *) 'Poissons ratio nu1 = ', nu1
       write(v, *) 'linear spring parameter alf = ', alf1
       write(v, *) 'angular spring parameter bet = ', bet1
       write(v, *) 'Young modulus Em1 of component solid 2 = ', Em2
       write(v, *) 'Poissons ratio nu2 = ', nu2
       write(v, *) 'linear spring parameter alf = ', alf2
       write(v, *) 'angular spring parameter bet = ', bet2
    end if
    write(v, *) 'solid medium density rs = ', rs
    write(v, *) 'integration time step dt = ', dt
    write(v, *) 'lattice unit step dl = ', dl
    write(v, *) 'Computational domain properties'
    write(v, *) 'x unit cell size ncx = ', gdim(1)
    write(v, *) 'y unit cell size ncy = ', gdim(2)
    write(v, *) 'z unit cell size ncz = ', gdim(3)
    write(v, *) 'number of solid points ns = ', nref
    write(v, *) 'LSM algorithm parameters'
    write(v, *) 'simulation duration itm = ', itm
    write(v, *) 'save parameter its = ', its
    write(v, *) 'nbo = ', nbo
    write(v, *) 'nna = ', nna
    write(v, *) 'layers = ', (layers(l), l = 1, 2 * nthread)
    write(v, *) 'layb = ', (layb(l), l = 1, 2 * nthread + 1)
    write(v, *) 'lib = ', (lib(l), l = 1, 2 * nthread + 1)
    write(v, *) 'lia = ', (lia(l), l = 1, 2 * nthread + 1)
    close(v)
  end subroutine write_settings

  ! The field files after a block of Verlet steps (XX_P.f:1546-1592).
  subroutine write_fields(iti_, itj_)
    integer, intent(in) :: iti_, itj_
    integer :: uu(4), v, ii, jj, kk, l
    real(8) :: fto(3), vto(3), rto(3), z3(3)
    z3 = 0.d0
    open(newunit=uu(1), file='LSM_coord_field', status='replace')
    open(newunit=uu(2), file='LSM_displ_field', status='replace')
    open(newunit=uu(3), file='LSM_velocity_field', status='replace')
    open(newunit=uu(4), file='LSM_force_field', status='replace')
    do kk = 1, gdim(3)
       do jj = 1, gdim(2)
          do ii = 1, gdim(1)
             l = mijks(ii, jj, kk)
             if (l /= 0) then
                write(uu(1), *) rr(:, l)
                write(uu(2), *) rr(:, l) - icors(:, l)
                write(uu(3), *) vv(:, l)
                write(uu(4), *) aa(:, l) * rs
             else
                write(uu(3), *) z3
                write(uu(4), *) z3
                write(uu(1), *) [ii, jj, kk]
                write(uu(2), *) z3
             end if
          end do
       end do
    end do
    do l = 1, 4
       close(uu(l))
    end do
    fto = 0.d0; vto = 0.d0; rto = 0.d0
    do l = 1, int(nref)
       fto = fto + aa(:, l) * rs
       vto = vto + vv(:, l)
       rto = rto + rr(:, l)
    end do
    open(newunit=v, file='write', status='replace')
    write(v, *) 'iti = ', iti_, 'itj = ', itj_
    write(v, *) 'FTO = ', fto
    write(v, *) 'VTO = ', vto
    write(v, *) 'RTO = ', rto
    close(v)
  end subroutine write_fields

  subroutine write_block(iti_, its_)
    integer, intent(in) :: iti_, its_
    integer :: v, pd, dd
    real(8) :: rncub
    character(len=3), parameter :: ax(3) = ['x  ', 'y  ', 'z  ']
    rncub = real(counts(2), 8)
    open(newunit=v, file='stress_tensor', position='append')
    write(v, *) 'Elastic medium properties'
    write(v, *) 'Ypung modulus Em = ', Em1, Em2
    write(v, *) 'Poisson ratio nu = ', nu1, nu2
    write(v, *) 'Simulation parameters'
    write(v, *) 'Cell volume vcell = ', real(ncx, 8) * ncy * ncz
    write(v, *) 'Solid part volume(elementary cubes) rncub = ', rncub
    write(v, *) 'Global strain exx0 = ', 2.d0 * gstrain
    if (its_ > 0) write(v, *) 'Verlet steps = ', iti_ * its_
    write(v, *) 'Simulation results'
    write(v, *) 'averaged stress tensor over ALL the cell'
    write(v, *) 'asigmap = '
    write(v, *) asig / max(rncub, 1.d0)
    write(v, *) 'Effective stiffness coefficients'
    if (lc <= 2) then
       dd = lc + 1
       write(v, *) 'E' // trim(ax(dd)) // trim(ax(dd)) // trim(ax(dd)) // trim(ax(dd)) // ' = ', &
            moduli(dd, dd)
       do pd = 1, 3
          if (pd == dd) cycle
          if (lc == 0 .and. pd == 3) then
             ! XX_P.f:1966 labels asigmap(3,3) of the xx case 'Eyyzz'
             write(v, *) 'Eyyzz = ', moduli(pd, pd)
          else
             write(v, *) 'E' // trim(ax(pd)) // trim(ax(pd)) // trim(ax(dd)) // trim(ax(dd)) // ' = ', &
                  moduli(pd, pd)
          end if
       end do
    else
       select case (lc)
       case (3)
          pd = 2; dd = 3
       case (4)
          pd = 3; dd = 1
       case default
          pd = 1; dd = 2
       end select
       write(v, *) 'E' // trim(ax(pd)) // trim(ax(dd)) // trim(ax(pd)) // trim(ax(dd)) // ' = ', &
            moduli(pd, dd)
    end if
    close(v)
  end subroutine write_block
end program psx_lsm_drv
