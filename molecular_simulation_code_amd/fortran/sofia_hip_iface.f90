! ISO_C_BINDING shim between the reference's Fortran host and libsofia_hip.so.
!
! Module sofia_hip_iface declares the C ABI (include/sofia_hip.h).  Module sofia_hip_dropin
! re-exposes the hot-path subroutines under the reference's names with a `_hip` suffix and
! the reference's argument lists; each one reads and writes `module global` exactly like
! the routine it replaces:
!
!   neighbor_wrapper(step)                  Source/mod_neighbor.f90:8-86
!   check_neighbor(nflag, step)             Source/mod_neighbor.f90:895-913
!   lj_cut_coul_dsf_nonewton_NOMIC(step)    Source/mod_force_nonbond.f90:105-196
!   lj_cut_coul_long_nonewton_NOMIC         Source/mod_force_nonbond.f90:300-400
!   correct_12/13/14_bonds                  Source/mod_force_correction.f90:6-278
!   nve_v(dt_partial), nve_x                Source/mod_ensemble.f90:11-87
!   pbc(step)                               Source/mod_pbc.f90:6-39
!   shake(iflag)                            Source/mod_shake.f90:8-85
!   dump_global_velocity_id_mpi(step, num)  Source/mod_datadump.f90:113-151
!   read_dump_global_id_v(filename)         Source/mod_datadump.f90:212-252
!
! A maintainer swaps the calls (or renames the wrappers) in force_driver / the time loops;
! sofia_hip_init is called once after the reference's initialisation sequence
! (Source/mod_restart_datadump.f90:139-233), when global holds the topology.
module sofia_hip_iface
  use iso_c_binding
  implicit none

  type, bind(c) :: sofia_params_t
     integer(c_int32_t) :: ntypes
     real(c_float) :: rcut, cut_coul, pad, coulpre, dt, kcal2amu
     real(c_double) :: rcut2, cut_coulsq, rv2, rvrc2, alpha, gewald, e_shift, f_shift
     real(c_double) :: ewald_p, ewald_f, a1, a2, a3, a4, a5, my_pis, my_pis_inv, amu2e
     real(c_double) :: fudge_12, fudge_13, fudge_14_coul, shake_tol
     integer(c_int32_t) :: shake_maxiter, precision, coul_flag, neigh_alloc0
  end type sofia_params_t

  interface
     integer(c_int) function sofia_ctx_create(p, np, box, ctx) bind(c, name="sofia_ctx_create")
       import :: c_int, c_int32_t, c_float, c_ptr, sofia_params_t
       type(sofia_params_t), intent(in) :: p
       integer(c_int32_t), value :: np
       real(c_float), value :: box
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int) function sofia_ctx_destroy(ctx) bind(c, name="sofia_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function sofia_set_forcefield(ctx, lj1, lj2, lj3, lj4, l141, l142, l143, l144, &
          mass) bind(c, name="sofia_set_forcefield")
       import :: c_int, c_ptr, c_float, c_double
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: lj1(*), lj2(*), lj3(*), lj4(*)
       real(c_double), intent(in) :: l141(*), l142(*), l143(*), l144(*), mass(*)
     end function
     integer(c_int) function sofia_set_exclusions(ctx, n1, n2, n3, spec) &
          bind(c, name="sofia_set_exclusions")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: n1(*), n2(*), n3(*), spec(*)
     end function
     integer(c_int) function sofia_set_bonded(ctx, bl, nb, nbr, al, na, nar, dl, nd, ndr) &
          bind(c, name="sofia_set_bonded")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx, nbr, nar, ndr
       integer(c_int32_t), intent(in) :: bl(*), al(*), dl(*)
       integer(c_int32_t), value :: nb, na, nd
     end function
     integer(c_int) function sofia_set_shake(ctx, ns, nsr, num, atoms, bond) &
          bind(c, name="sofia_set_shake")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, nsr
       integer(c_int32_t), value :: ns
       integer(c_int32_t), intent(in) :: num(*), atoms(*)
       real(c_double), intent(in) :: bond(*)
     end function
     ! position is type(ainfo) and v/ff are type(coords): assumed-type dummies pass them
     ! without copies (same 16 B / 12 B layouts as sofia_ainfo / sofia_coords)
     integer(c_int) function sofia_put_state(ctx, pos, q, v, gid) bind(c, name="sofia_put_state")
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       type(*), intent(in) :: pos(*), v(*)
       real(c_float), intent(in) :: q(*)
       integer(c_int32_t), intent(in) :: gid(*)
     end function
     integer(c_int) function sofia_get_state(ctx, pos, v, gid) bind(c, name="sofia_get_state")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       type(*), intent(inout) :: pos(*), v(*)
       integer(c_int32_t), intent(inout) :: gid(*)
     end function
     integer(c_int) function sofia_put_displacement(ctx, dx, dy, dz) &
          bind(c, name="sofia_put_displacement")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: dx(*), dy(*), dz(*)
     end function
     integer(c_int) function sofia_get_displacement(ctx, dx, dy, dz) &
          bind(c, name="sofia_get_displacement")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(inout) :: dx(*), dy(*), dz(*)
     end function
     ! The same two C functions with only some arrays (the others NULL): a wrapper moves
     ! just the arrays its routine reads (put) or writes (get).
     integer(c_int) function sofia_put_pos(ctx, pos, q, v, gid) bind(c, name="sofia_put_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, q, v, gid
       type(*), intent(in) :: pos(*)
     end function
     integer(c_int) function sofia_put_pos_v(ctx, pos, q, v, gid) bind(c, name="sofia_put_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, q, gid
       type(*), intent(in) :: pos(*), v(*)
     end function
     integer(c_int) function sofia_put_v(ctx, pos, q, v, gid) bind(c, name="sofia_put_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, pos, q, gid
       type(*), intent(in) :: v(*)
     end function
     integer(c_int) function sofia_get_pos(ctx, pos, v, gid) bind(c, name="sofia_get_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, v, gid
       type(*), intent(inout) :: pos(*)
     end function
     integer(c_int) function sofia_get_v(ctx, pos, v, gid) bind(c, name="sofia_get_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, pos, gid
       type(*), intent(inout) :: v(*)
     end function
     ! page-locked host memory for position / v / ff (bind with c_f_pointer)
     integer(c_int) function sofia_host_alloc(bytes, p) bind(c, name="sofia_host_alloc")
       import :: c_int, c_ptr, c_size_t
       integer(c_size_t), value :: bytes
       type(c_ptr), intent(out) :: p
     end function
     integer(c_int) function sofia_host_free(p) bind(c, name="sofia_host_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: p
     end function
     integer(c_int) function sofia_neighbor_wrapper(ctx, step, tt) bind(c, name="sofia_neighbor_wrapper")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: step
       integer(c_int32_t), intent(inout) :: tt(*)
     end function
     integer(c_int) function sofia_rdf_calc(ctx, itype, jtype, delta, nbins, rdf, numob) &
          bind(c, name="sofia_rdf_calc")
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: itype, jtype, nbins
       real(c_float), value :: delta
       real(c_float), intent(inout) :: rdf(*)
       integer(c_int32_t), intent(inout) :: numob
     end function
     integer(c_int) function sofia_get_topology(ctx, q, n1, n2, n3, spec, bl, al, dl, il, sa) &
          bind(c, name="sofia_get_topology")
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(inout) :: q(*)
       integer(c_int32_t), intent(inout) :: n1(*), n2(*), n3(*), spec(*), bl(*), al(*), dl(*)
       type(c_ptr), value :: il, sa
     end function
     integer(c_int) function sofia_get_nlist(ctx, nlist, numneigh, na) bind(c, name="sofia_get_nlist")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(inout) :: nlist(*), numneigh(*)
       integer(c_int32_t), intent(out) :: na
     end function
     integer(c_int) function sofia_check_neighbor(ctx, nflag) bind(c, name="sofia_check_neighbor")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: nflag
     end function
     integer(c_int) function sofia_lj_cut_coul_dsf(ctx, ff, ff64, pot, ecoul, vir) &
          bind(c, name="sofia_lj_cut_coul_dsf")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, ff64
       type(*), intent(inout) :: ff(*)
       real(c_double), intent(inout) :: pot, ecoul, vir(3)
     end function
     integer(c_int) function sofia_lj_cut_coul_dsf_async(ctx, mixed) &
          bind(c, name="sofia_lj_cut_coul_dsf_async")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mixed
     end function
     integer(c_int) function sofia_nonbond_wait(ctx, ff, pot, ecoul, vir) &
          bind(c, name="sofia_nonbond_wait")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       type(*), intent(inout) :: ff(*)
       real(c_double), intent(inout) :: pot, ecoul, vir(3)
     end function
     integer(c_int) function sofia_lj_cut_coul_long(ctx, ff, pot, ecoul, vir) &
          bind(c, name="sofia_lj_cut_coul_long")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       type(*), intent(inout) :: ff(*)
       real(c_double), intent(inout) :: pot, ecoul, vir(3)
     end function
     integer(c_int) function sofia_correct_bonds(ctx, which, ff, ec, vir) bind(c, name="sofia_correct_bonds")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which
       type(*), intent(inout) :: ff(*)
       real(c_double), intent(inout) :: ec, vir(3)
     end function
     integer(c_int) function sofia_nve_v(ctx, dtp, ff, ke) bind(c, name="sofia_nve_v")
       import :: c_int, c_ptr, c_float, c_double
       type(c_ptr), value :: ctx
       real(c_float), value :: dtp
       type(*), intent(in) :: ff(*)
       real(c_double), intent(out) :: ke
     end function
     integer(c_int) function sofia_nve_x(ctx) bind(c, name="sofia_nve_x")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function sofia_pbc(ctx) bind(c, name="sofia_pbc")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function sofia_shake(ctx, iflag, ff, vir, maxit) bind(c, name="sofia_shake")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: iflag
       type(*), intent(inout) :: ff(*)
       real(c_double), intent(inout) :: vir(3)
       integer(c_int32_t), intent(out) :: maxit
     end function
     integer(c_int) function sofia_write_restart(path, box, temp, num, timeclock, np, gid, pos, v) &
          bind(c, name="sofia_write_restart")
       import :: c_int, c_char, c_float, c_int32_t
       character(kind=c_char), intent(in) :: path(*)
       real(c_float), value :: box, temp, timeclock
       integer(c_int32_t), value :: num, np
       integer(c_int32_t), intent(in) :: gid(*)
       real(c_float), intent(in) :: pos(*), v(*)
     end function
     integer(c_int) function sofia_read_restart(path, np, box, temp, num, timeclock, pos, v) &
          bind(c, name="sofia_read_restart")
       import :: c_int, c_char, c_float, c_int32_t
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: np
       real(c_float), intent(out) :: box, temp, timeclock
       integer(c_int32_t), intent(out) :: num
       real(c_float), intent(out) :: pos(*), v(*)
     end function
     integer(c_int) function sofia_last_error(buf, n) bind(c, name="sofia_last_error")
       import :: c_int, c_char, c_size_t
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_size_t), value :: n
     end function
  end interface

contains

  subroutine sofia_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    integer(c_int) :: dummy
    integer :: k
    character(len=512) :: msg
    if (rc == 0) return
    dummy = sofia_last_error(buf, int(512, c_size_t))
    msg = ' '
    do k = 1, 512
       if (buf(k) == c_null_char) exit
       msg(k:k) = buf(k)
    end do
    print *, 'libsofia_hip error in ', where, ': ', trim(msg)
    stop
  end subroutine sofia_check

end module sofia_hip_iface


module sofia_hip_dropin
  use iso_c_binding
  use global
  use sofia_hip_iface
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  integer(c_int32_t), allocatable, save :: tt_buf(:)

contains

  ! Create the device context from module global once the reference has read its inputs
  ! (constants as in Source/get_started.f90:25-117).
  subroutine sofia_hip_init(precision)
    integer, intent(in) :: precision
    type(sofia_params_t) :: p
    real(c_double), allocatable :: l14(:,:)
    integer :: a, b, nt
    nt = numAtomType
    p%ntypes = nt; p%rcut = rcut; p%cut_coul = cut_coul; p%pad = neighbor_pad
    p%coulpre = coulpre; p%dt = dt; p%kcal2amu = kcal2amu
    p%rcut2 = rcut2; p%cut_coulsq = cut_coulsq; p%rv2 = rv2; p%rvrc2 = rvrc2
    p%alpha = alpha; p%gewald = gewald; p%e_shift = e_shift; p%f_shift = f_shift
    p%ewald_p = EWALD_P; p%ewald_f = EWALD_F
    p%a1 = A1; p%a2 = A2; p%a3 = A3; p%a4 = A4; p%a5 = A5
    p%my_pis = MY_PIS; p%my_pis_inv = MY_PIS_INV; p%amu2e = amu2e
    p%fudge_12 = fudge_12; p%fudge_13 = fudge_13; p%fudge_14_coul = fudge_14_coul
    p%shake_tol = tolerance; p%shake_maxiter = maxiter
    p%precision = precision; p%coul_flag = coul_flag; p%neigh_alloc0 = neigh_alloc
    call sofia_check(sofia_ctx_create(p, int(np, c_int32_t), real(box, c_float), ctx), 'ctx_create')
    ! lj14 tables are (itype, jtype); the ABI wants row-major (itype-1)*nt + jtype
    allocate(l14(nt*nt, 4))
    do a = 1, nt
       do b = 1, nt
          l14((a-1)*nt + b, 1) = lj14_1(a, b); l14((a-1)*nt + b, 2) = lj14_2(a, b)
          l14((a-1)*nt + b, 3) = lj14_3(a, b); l14((a-1)*nt + b, 4) = lj14_4(a, b)
       end do
    end do
    call sofia_check(sofia_set_forcefield(ctx, lj1, lj2, lj3, lj4, l14(:,1), l14(:,2), l14(:,3), &
         l14(:,4), mass), 'set_forcefield')
    call sofia_check(sofia_set_exclusions(ctx, num1bond, num2bond, num3bond, specbond), 'set_exclusions')
    call sofia_check(sofia_set_bonded(ctx, bondlist, int(numbond, c_int32_t), c_null_ptr, anglelist, &
         int(numangle, c_int32_t), c_null_ptr, dihedlist, int(numdihed, c_int32_t), c_null_ptr), &
         'set_bonded')
    allocate(tt_buf(np))
  end subroutine sofia_hip_init

  ! neighbor_wrapper, Source/mod_neighbor.f90:8-86.  The device sorts and lists, and returns
  ! the arrays it permuted itself in the host's layout (sofia_get_topology: q, num1/2/3bond,
  ! specbond, the bond, angle and dihedral lists that sofia_hip_init gave it).  The arrays
  ! the device does not hold (mol, integrate_flag, improlist, the shakeatom records) are
  ! re-indexed through tt here, as sort_posit_charge, sort_improlist and sort_shakeatom do
  ! (Source/mod_neighbor.f90:946-1005, 1043-1061, 1126-1168).
  subroutine neighbor_wrapper_hip(step)
    integer :: step, i
    nn_builds = nn_builds + 1
    call sofia_check(sofia_put_state(ctx, position, q, v, global_id), 'put_state')
    call sofia_check(sofia_neighbor_wrapper(ctx, int(step, c_int32_t), tt_buf), 'neighbor_wrapper')
    call sofia_check(sofia_get_state(ctx, position, v, global_id), 'get_state')
    call sofia_check(sofia_get_topology(ctx, q, num1bond, num2bond, num3bond, specbond, bondlist, &
         anglelist, dihedlist, c_null_ptr, c_null_ptr), 'get_topology')
    do i = 1, np
       molsort(:, tt_buf(i)) = mol(:, i)
       integrate_flag_ss(tt_buf(i)) = integrate_flag(i)
    end do
    mol(:,:) = molsort(:,:)
    integrate_flag(:) = integrate_flag_ss(:)
    do i = 1, numimpro
       improlist(1:4, i) = tt_buf(improlist(1:4, i))
    end do
    do i = 1, nshake
       select case (shakeatom(i)%num)
       case (1)
          shakeatom(i)%atom1 = tt_buf(shakeatom(i)%atom1)
          shakeatom(i)%atom2 = tt_buf(shakeatom(i)%atom2)
       case (3, -3)
          shakeatom(i)%atom1 = tt_buf(shakeatom(i)%atom1)
          shakeatom(i)%atom2 = tt_buf(shakeatom(i)%atom2)
          shakeatom(i)%atom3 = tt_buf(shakeatom(i)%atom3)
       case (4)
          shakeatom(i)%atom1 = tt_buf(shakeatom(i)%atom1)
          shakeatom(i)%atom2 = tt_buf(shakeatom(i)%atom2)
          shakeatom(i)%atom3 = tt_buf(shakeatom(i)%atom3)
          shakeatom(i)%atom4 = tt_buf(shakeatom(i)%atom4)
       end select
    end do
    drx(:) = 0.0; dry(:) = 0.0; drz(:) = 0.0
  end subroutine neighbor_wrapper_hip

  ! RDF_calc, Source/mod_force_nonbond.f90:708-775 (type 3 with type 3, as the reference),
  ! on the list of the last neighbor_wrapper_hip: the k-th element of rdf receives bin k
  ! (node k, from 0); bins past the end of rdf are not counted (`size` is a variable of
  ! module global, so the extent comes from the bounds).
  subroutine RDF_calc_hip
    integer(c_int32_t) :: nob
    nob = int(numob_rdf, c_int32_t)
    call sofia_check(sofia_put_pos(ctx, position, c_null_ptr, c_null_ptr, c_null_ptr), 'put_state')
    call sofia_check(sofia_rdf_calc(ctx, 3_c_int32_t, 3_c_int32_t, rdf_delta, &
         int(ubound(rdf, 1) - lbound(rdf, 1) + 1, c_int32_t), rdf, nob), 'rdf_calc')
    numob_rdf = nob
  end subroutine RDF_calc_hip

  subroutine check_neighbor_hip(nflag, step)
    integer :: nflag, step
    integer(c_int32_t) :: f
    call sofia_check(sofia_put_displacement(ctx, drx, dry, drz), 'put_displacement')
    call sofia_check(sofia_check_neighbor(ctx, f), 'check_neighbor')
    nflag = f
  end subroutine check_neighbor_hip

  ! Each wrapper copies in the arrays its routine reads and copies out the arrays it
  ! writes; q, global_id and the topology change only in neighbor_wrapper_hip.
  subroutine lj_cut_coul_dsf_nonewton_NOMIC_hip(step)
    integer :: step
    real(c_double) :: vir(3)
    call sofia_check(sofia_put_pos(ctx, position, c_null_ptr, c_null_ptr, c_null_ptr), 'put_state')
    call sofia_check(sofia_lj_cut_coul_dsf(ctx, ff, c_null_ptr, potential, e_coul, vir), 'dsf')
  end subroutine lj_cut_coul_dsf_nonewton_NOMIC_hip

  ! The offload pattern of Source/mod_force.f90:66-156: start the pair kernel, do host work,
  ! then wait (offload ... signal(sig) / offload_wait).
  subroutine lj_cut_coul_dsf_nonewton_MIC_signal_hip(step)
    integer :: step
    call sofia_check(sofia_put_pos(ctx, position, c_null_ptr, c_null_ptr, c_null_ptr), 'put_state')
    call sofia_check(sofia_lj_cut_coul_dsf_async(ctx, 0_c_int32_t), 'dsf_async')
  end subroutine lj_cut_coul_dsf_nonewton_MIC_signal_hip

  subroutine lj_cut_coul_dsf_nonewton_MIC_wait_hip
    real(c_double) :: vir(3)
    call sofia_check(sofia_nonbond_wait(ctx, ff, potential, e_coul, vir), 'nonbond_wait')
  end subroutine lj_cut_coul_dsf_nonewton_MIC_wait_hip

  subroutine lj_cut_coul_long_nonewton_NOMIC_hip
    real(c_double) :: vir(3)
    call sofia_check(sofia_put_pos(ctx, position, c_null_ptr, c_null_ptr, c_null_ptr), 'put_state')
    call sofia_check(sofia_lj_cut_coul_long(ctx, ff, potential, e_coul, vir), 'coul_long')
    virialx = vir(1); virialy = vir(2); virialz = vir(3)
  end subroutine lj_cut_coul_long_nonewton_NOMIC_hip

  subroutine correct_12_bonds_hip
    real(c_double) :: vir(3)
    vir = (/ v_bondx, v_bondy, v_bondz /)
    call sofia_check(sofia_correct_bonds(ctx, 12_c_int32_t, ff_bond, e_coul_bond, vir), 'correct_12')
    v_bondx = vir(1); v_bondy = vir(2); v_bondz = vir(3)
  end subroutine correct_12_bonds_hip

  subroutine correct_13_bonds_hip
    real(c_double) :: vir(3)
    vir = (/ v_anglex, v_angley, v_anglez /)
    call sofia_check(sofia_correct_bonds(ctx, 13_c_int32_t, ff_angle, e_coul_angle, vir), 'correct_13')
    v_anglex = vir(1); v_angley = vir(2); v_anglez = vir(3)
  end subroutine correct_13_bonds_hip

  subroutine correct_14_bonds_hip
    real(c_double) :: vir(3)
    vir = (/ v_dihedx, v_dihedy, v_dihedz /)
    call sofia_check(sofia_correct_bonds(ctx, 14_c_int32_t, ff_dihed, e_coul_dihed, vir), 'correct_14')
    v_dihedx = vir(1); v_dihedy = vir(2); v_dihedz = vir(3)
  end subroutine correct_14_bonds_hip

  subroutine nve_v_hip(dt_partial)
    real*4 :: dt_partial
    call sofia_check(sofia_put_v(ctx, c_null_ptr, c_null_ptr, v, c_null_ptr), 'put_state')
    call sofia_check(sofia_nve_v(ctx, dt_partial, ff, ke), 'nve_v')
    call sofia_check(sofia_get_v(ctx, c_null_ptr, v, c_null_ptr), 'get_state')
  end subroutine nve_v_hip

  subroutine nve_x_hip
    call sofia_check(sofia_put_pos_v(ctx, position, c_null_ptr, v, c_null_ptr), 'put_state')
    call sofia_check(sofia_put_displacement(ctx, drx, dry, drz), 'put_displacement')
    call sofia_check(sofia_nve_x(ctx), 'nve_x')
    call sofia_check(sofia_get_pos(ctx, position, c_null_ptr, c_null_ptr), 'get_state')
    call sofia_check(sofia_get_displacement(ctx, drx, dry, drz), 'get_displacement')
  end subroutine nve_x_hip

  subroutine pbc_hip(step)
    integer :: step
    call sofia_check(sofia_put_pos(ctx, position, c_null_ptr, c_null_ptr, c_null_ptr), 'put_state')
    call sofia_check(sofia_pbc(ctx), 'pbc')
    call sofia_check(sofia_get_pos(ctx, position, c_null_ptr, c_null_ptr), 'get_state')
  end subroutine pbc_hip

  subroutine shake_hip(iflag)
    integer :: iflag
    real(c_double) :: vir(3)
    integer(c_int32_t) :: its
    vir = (/ virialx, virialy, virialz /)
    call sofia_check(sofia_put_pos_v(ctx, position, c_null_ptr, v, c_null_ptr), 'put_state')
    call sofia_check(sofia_shake(ctx, int(iflag, c_int32_t), ff, vir, its), 'shake')
    virialx = vir(1); virialy = vir(2); virialz = vir(3)
  end subroutine shake_hip

  ! restart_<step> under restart_files_path, as dump_global_velocity_id_mpi writes it (the
  ! reference also flushes its observable files first; that stays with the caller).
  subroutine dump_global_velocity_id_mpi_hip(step, num)
    integer, intent(in) :: step, num
    character(len=50) :: ints
    character(len=256) :: filename
    real(c_float), allocatable :: xyz(:, :), vel(:, :)
    integer(c_int32_t), allocatable :: gid(:)
    integer :: i
    allocate(xyz(3, np), vel(3, np), gid(np))
    do i = 1, np
       xyz(1, i) = position(i)%x; xyz(2, i) = position(i)%y; xyz(3, i) = position(i)%z
       vel(1, i) = v(i)%x; vel(2, i) = v(i)%y; vel(3, i) = v(i)%z
       gid(i) = global_id(i)
    end do
    write(ints, '(i8)') step
    filename = trim(restart_files_path)//'/'//'restart_'//trim(adjustl(ints))
    call sofia_check(sofia_write_restart(trim(filename)//c_null_char, box, temp, int(num, c_int32_t), &
                                         timeclock, int(np, c_int32_t), gid, xyz, vel), 'write_restart')
  end subroutine dump_global_velocity_id_mpi_hip

  ! read_dump_global_id_v: box, timeclock and the state placed by global id; hbox, ibox and
  ! vol follow the box as in the reference.
  subroutine read_dump_global_id_v_hip(filename)
    character(len=*), intent(in) :: filename
    real(c_float), allocatable :: xyz(:, :), vel(:, :)
    real(c_float) :: junk
    integer(c_int32_t) :: read_num
    integer :: i
    allocate(xyz(3, np), vel(3, np))
    call sofia_check(sofia_read_restart(trim(filename)//c_null_char, int(np, c_int32_t), box, junk, &
                                        read_num, timeclock, xyz, vel), 'read_restart')
    hbox = box / 2.0
    ibox = 1.0 / box
    vol = box * box * box
    do i = 1, np
       position(i)%x = xyz(1, i); position(i)%y = xyz(2, i); position(i)%z = xyz(3, i)
       v(i)%x = vel(1, i); v(i)%y = vel(2, i); v(i)%z = vel(3, i)
    end do
  end subroutine read_dump_global_id_v_hip

end module sofia_hip_dropin
