! Driver for the reference's own Fortran routines -- TEST INFRASTRUCTURE ONLY.
!
! oracle/build_ref.sh compiles this file together with the reference modules where they
! lie (/root/reference/Source: global, mod_minimg, mod_force_nonbond,
! mod_force_correction, mod_force_dihed, mod_force_bond, mod_force_angle, mod_pbc,
! mod_force_impro_amber, mod_force_dihed_amber) into
! oracle/_ref/libsofia_ref.so.  The entry points below only copy caller arrays into
! `module global`, call one reference subroutine unchanged, and copy its outputs back,
! so the golden-vector script in tests/golden/ can compare the C oracle with the
! reference itself.  Nothing here is part of the product.
module sofia_ref_driver
  use iso_c_binding
  use global
  use mod_force_nonbond
  use mod_force_correction
  use mod_force_dihed
  use mod_force_bond
  use mod_force_angle
  use mod_pbc
  use mod_force_impro_amber
  use mod_force_dihed_amber
  implicit none
contains

  ! Scalars of module global (precisions as declared in Source/global.f90).
  subroutine ref_set_scalars(c_box, c_ibox, c_rcut2, c_cut_coulsq, c_alpha, c_gewald, &
       c_e_shift, c_f_shift, c_coulpre, c_ewald_f, c_ewald_p, c_a, c_pis, c_pis_inv, &
       c_fudge12, c_fudge13, c_fudge14) bind(c, name="ref_set_scalars")
    real(c_float), value :: c_box, c_ibox, c_coulpre
    real(c_double), value :: c_rcut2, c_cut_coulsq, c_alpha, c_gewald, c_e_shift, c_f_shift
    real(c_double), value :: c_ewald_f, c_ewald_p, c_pis, c_pis_inv
    real(c_double), value :: c_fudge12, c_fudge13, c_fudge14
    real(c_double), intent(in) :: c_a(5)
    box = c_box; ibox = c_ibox; coulpre = c_coulpre
    rcut2 = c_rcut2; cut_coulsq = c_cut_coulsq
    alpha = c_alpha; gewald = c_gewald
    e_shift = c_e_shift; f_shift = c_f_shift
    EWALD_F = c_ewald_f; EWALD_P = c_ewald_p
    A1 = c_a(1); A2 = c_a(2); A3 = c_a(3); A4 = c_a(4); A5 = c_a(5)
    MY_PIS = c_pis; MY_PIS_INV = c_pis_inv
    fudge_12 = c_fudge12; fudge_13 = c_fudge13; fudge_14_coul = c_fudge14
  end subroutine ref_set_scalars

  ! Atoms (x, y, z, type), charges and the LJ tables.
  subroutine ref_set_atoms(n, nt, xyz, typ, qq, l1, l2, l3, l4, l141, l142, l143, l144) &
       bind(c, name="ref_set_atoms")
    integer(c_int), value :: n, nt
    real(c_float), intent(in) :: xyz(3, n), qq(n), l1(nt*nt), l2(nt*nt), l3(nt*nt), l4(nt*nt)
    integer(c_int), intent(in) :: typ(n)
    real(c_double), intent(in) :: l141(nt*nt), l142(nt*nt), l143(nt*nt), l144(nt*nt)
    integer :: i, a, b
    np = n; numAtomType = nt
    if (allocated(position)) deallocate(position, q, ff, ffh, ff_bond, ff_angle, ff_dihed)
    if (allocated(lj1)) deallocate(lj1, lj2, lj3, lj4, lj14_1, lj14_2, lj14_3, lj14_4)
    allocate(position(n), q(n), ff(n), ffh(n), ff_bond(n), ff_angle(n), ff_dihed(n))
    allocate(lj1(nt*nt), lj2(nt*nt), lj3(nt*nt), lj4(nt*nt))
    allocate(lj14_1(nt,nt), lj14_2(nt,nt), lj14_3(nt,nt), lj14_4(nt,nt))
    do i = 1, n
       position(i)%x = xyz(1, i); position(i)%y = xyz(2, i); position(i)%z = xyz(3, i)
       position(i)%type = typ(i)
       q(i) = qq(i)
    end do
    lj1 = l1; lj2 = l2; lj3 = l3; lj4 = l4
    ! lj14 tables are (itype, jtype) arrays; the flat input is row-major (i-1)*nt + j
    do a = 1, nt
       do b = 1, nt
          lj14_1(a, b) = l141((a-1)*nt + b); lj14_2(a, b) = l142((a-1)*nt + b)
          lj14_3(a, b) = l143((a-1)*nt + b); lj14_4(a, b) = l144((a-1)*nt + b)
       end do
    end do
  end subroutine ref_set_atoms

  subroutine get_forces(src, out)
    type(coords), intent(in) :: src(:)
    real(c_float), intent(out) :: out(3, *)
    integer :: i
    do i = 1, np
       out(1, i) = src(i)%x; out(2, i) = src(i)%y; out(3, i) = src(i)%z
    end do
  end subroutine get_forces

  subroutine zero_forces()
    integer :: i
    do i = 1, np
       ff(i) = coords(0.0, 0.0, 0.0); ffh(i) = coords(0.0, 0.0, 0.0)
       ff_bond(i) = coords(0.0, 0.0, 0.0); ff_angle(i) = coords(0.0, 0.0, 0.0)
       ff_dihed(i) = coords(0.0, 0.0, 0.0)
    end do
  end subroutine zero_forces

  ! lj_cut_coul_dsf_nonewton_NOMIC (Source/mod_force_nonbond.f90:105-196) on a given
  ! Verlet list (1-based ids, stride na).
  subroutine ref_dsf(na, nl, nn, fout, epot, ecoul) bind(c, name="ref_dsf")
    integer(c_int), value :: na
    integer(c_int), intent(in) :: nl(na*np), nn(np)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: epot, ecoul
    if (allocated(nlist)) deallocate(nlist, numneigh)
    allocate(nlist(na*np), numneigh(np))
    nlist = nl; numneigh = nn; neigh_alloc = na
    call zero_forces()
    potential = 0.0d0; e_coul = 0.0d0
    call lj_cut_coul_dsf_nonewton_NOMIC(0)
    call get_forces(ff, fout)
    epot = potential; ecoul = e_coul
  end subroutine ref_dsf

  ! RDF_calc (Source/mod_force_nonbond.f90:708-775) on a given Verlet list: rdf is
  ! allocated as rdf(0:nb-1) and zeroed, so out(k+1) is the count of bin k; numob_rdf in and
  ! out.  The caller sizes nb past sqrt(rcut2)/delta (the routine has no bound check).
  subroutine ref_rdf(na, nl, nn, delta, nb, out, nob) bind(c, name="ref_rdf")
    integer(c_int), value :: na, nb
    integer(c_int), intent(in) :: nl(na*np), nn(np)
    real(c_float), value :: delta
    real(c_float), intent(out) :: out(nb)
    integer(c_int), intent(inout) :: nob
    if (allocated(nlist)) deallocate(nlist, numneigh)
    allocate(nlist(na*np), numneigh(np))
    nlist = nl; numneigh = nn; neigh_alloc = na
    if (allocated(rdf)) deallocate(rdf)
    allocate(rdf(0:nb-1))
    rdf = 0.0
    rdf_delta = delta
    numob_rdf = nob
    call RDF_calc
    out(1:nb) = rdf(0:nb-1)
    nob = numob_rdf
  end subroutine ref_rdf

  ! lj_cut_coul_long_nonewton_NOMIC (Source/mod_force_nonbond.f90:300-400).
  subroutine ref_coul_long(na, nl, nn, fout, epot, ecoul, vir) bind(c, name="ref_coul_long")
    integer(c_int), value :: na
    integer(c_int), intent(in) :: nl(na*np), nn(np)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: epot, ecoul, vir(3)
    if (allocated(nlist)) deallocate(nlist, numneigh)
    allocate(nlist(na*np), numneigh(np))
    nlist = nl; numneigh = nn; neigh_alloc = na
    call zero_forces()
    potential = 0.0d0; e_coul = 0.0d0
    call lj_cut_coul_long_nonewton_NOMIC
    call get_forces(ff, fout)
    epot = potential; ecoul = e_coul
    vir(1) = virialx; vir(2) = virialy; vir(3) = virialz
  end subroutine ref_coul_long

  ! lj_cut_coul_dsf_N2 (Source/mod_force_nonbond.f90:409-502) with the exclusion table.
  subroutine ref_dsf_n2(n1, n2, n3, spec, fout, epot, ecoul) bind(c, name="ref_dsf_n2")
    integer(c_int), intent(in) :: n1(np), n2(np), n3(np), spec(12, np)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: epot, ecoul
    if (allocated(num1bond)) deallocate(num1bond, num2bond, num3bond, specbond)
    allocate(num1bond(np), num2bond(np), num3bond(np), specbond(12, np))
    num1bond = n1; num2bond = n2; num3bond = n3; specbond = spec
    call zero_forces()
    potential = 0.0d0; e_coul = 0.0d0
    call lj_cut_coul_dsf_N2(0)
    call get_forces(ff, fout)
    epot = potential; ecoul = e_coul
  end subroutine ref_dsf_n2

  ! correct_12_bonds / correct_13_bonds / correct_14_bonds
  ! (Source/mod_force_correction.f90:6-278).  kind = 12, 13 or 14; lst holds the
  ! reference list (bondlist(3,n) / anglelist(4,n) / dihedlist(5,n)).
  subroutine ref_correct(kind, n, width, lst, fout, ec, vir) bind(c, name="ref_correct")
    integer(c_int), value :: kind, n, width
    integer(c_int), intent(in) :: lst(width, n)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: ec, vir(3)
    call zero_forces()
    e_coul_bond = 0.0d0; e_coul_angle = 0.0d0; e_coul_dihed = 0.0d0
    v_bondx = 0.0d0; v_bondy = 0.0d0; v_bondz = 0.0d0
    v_anglex = 0.0d0; v_angley = 0.0d0; v_anglez = 0.0d0
    v_dihedx = 0.0d0; v_dihedy = 0.0d0; v_dihedz = 0.0d0
    if (kind == 12) then
       if (allocated(bondlist)) deallocate(bondlist)
       allocate(bondlist(3, n)); bondlist = lst; numbond = n
       call correct_12_bonds
       call get_forces(ff_bond, fout); ec = e_coul_bond
       vir(1) = v_bondx; vir(2) = v_bondy; vir(3) = v_bondz
    else if (kind == 13) then
       if (allocated(anglelist)) deallocate(anglelist)
       allocate(anglelist(4, n)); anglelist = lst; numangle = n
       call correct_13_bonds
       call get_forces(ff_angle, fout); ec = e_coul_angle
       vir(1) = v_anglex; vir(2) = v_angley; vir(3) = v_anglez
    else
       if (allocated(dihedlist)) deallocate(dihedlist)
       allocate(dihedlist(5, n)); dihedlist = lst; numdihed = n
       call correct_14_bonds
       call get_forces(ff_dihed, fout); ec = e_coul_dihed
       vir(1) = v_dihedx; vir(2) = v_dihedy; vir(3) = v_dihedz
    end if
  end subroutine ref_correct

  ! dihedral_opls_DSF (Source/mod_force_dihed.f90:816-1092) with every torsion
  ! coefficient zero, which leaves exactly the scaled 1-4 DSF + LJ pair term
  ! (lines 1003-1048) in ffh, e_coul_14 and pot_14.
  subroutine ref_pair14(n, lst, fout, ec14, pe14) bind(c, name="ref_pair14")
    integer(c_int), value :: n
    integer(c_int), intent(in) :: lst(5, n)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: ec14, pe14
    integer :: ntype
    if (allocated(dihedlist)) deallocate(dihedlist)
    allocate(dihedlist(5, n)); dihedlist = lst; numdihed = n
    ntype = maxval(lst(5, :))
    if (allocated(dihedcoeff)) deallocate(dihedcoeff)
    allocate(dihedcoeff(4, ntype)); dihedcoeff = 0.0d0
    call zero_forces()
    e_coul_14 = 0.0d0; pot_14 = 0.0d0
    call dihedral_opls_DSF(0)
    call get_forces(ffh, fout)
    ec14 = e_coul_14; pe14 = pot_14
  end subroutine ref_pair14

  ! bond_harmonic / angle_harmonic (Source/mod_force_bond.f90:6-75,
  ! Source/mod_force_angle.f90:10-122).
  subroutine ref_bond(n, lst, nbt, kb, r0, fout, eb) bind(c, name="ref_bond")
    integer(c_int), value :: n, nbt
    integer(c_int), intent(in) :: lst(3, n)
    real(c_double), intent(in) :: kb(nbt), r0(nbt)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: eb
    if (allocated(bondlist)) deallocate(bondlist)
    allocate(bondlist(3, n)); bondlist = lst; numbond = n
    if (allocated(bondcoeff)) deallocate(bondcoeff)
    allocate(bondcoeff(2, nbt)); bondcoeff(1, :) = kb; bondcoeff(2, :) = r0
    call zero_forces()
    e_bond = 0.0d0
    call bond_harmonic(0)
    call get_forces(ff_bond, fout)
    eb = e_bond
  end subroutine ref_bond

  subroutine ref_angle(n, lst, nat, k, th0, fout, ea) bind(c, name="ref_angle")
    integer(c_int), value :: n, nat
    integer(c_int), intent(in) :: lst(4, n)
    real(c_double), intent(in) :: k(nat), th0(nat)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: ea
    if (allocated(anglelist)) deallocate(anglelist)
    allocate(anglelist(4, n)); anglelist = lst; numangle = n
    if (allocated(anglecoeff)) deallocate(anglecoeff)
    allocate(anglecoeff(2, nat)); anglecoeff(1, :) = k; anglecoeff(2, :) = th0
    call zero_forces()
    e_angle = 0.0d0
    call angle_harmonic(0)
    call get_forces(ff_angle, fout)
    ea = e_angle
  end subroutine ref_angle

  ! improper_harmonic_lammps14 (Source/mod_force_dihed.f90:1546-1724).
  subroutine ref_improper(n, lst, nit, k, chi0, fout, ei, vir) bind(c, name="ref_improper")
    integer(c_int), value :: n, nit
    integer(c_int), intent(in) :: lst(5, n)
    real(c_double), intent(in) :: k(nit), chi0(nit)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: ei, vir(3)
    integer :: i
    if (allocated(improlist)) deallocate(improlist)
    allocate(improlist(5, n)); improlist = lst; numimpro = n
    if (allocated(improcoeff)) deallocate(improcoeff)
    allocate(improcoeff(2, nit)); improcoeff(1, :) = k; improcoeff(2, :) = chi0
    if (allocated(ff_impro)) deallocate(ff_impro)
    allocate(ff_impro(np))
    do i = 1, np
       ff_impro(i) = coords(0.0, 0.0, 0.0)
    end do
    e_improper = 0.0d0
    v_improx = 0.0d0; v_improy = 0.0d0; v_improz = 0.0d0
    call improper_harmonic_lammps14(0)
    call get_forces(ff_impro, fout)
    ei = e_improper
    vir(1) = v_improx; vir(2) = v_improy; vir(3) = v_improz
  end subroutine ref_improper

  ! impro_amber_force (Source/mod_force_impro_amber.f90:8-230).
  subroutine ref_improper_cvff(n, lst, nit, k, sgn, mlt, fout, ei, vir) &
       bind(c, name="ref_improper_cvff")
    integer(c_int), value :: n, nit
    integer(c_int), intent(in) :: lst(5, n), sgn(nit), mlt(nit)
    real(c_double), intent(in) :: k(nit)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: ei, vir(3)
    integer :: i
    if (allocated(improlist)) deallocate(improlist)
    allocate(improlist(5, n)); improlist = lst; numimpro = n
    if (allocated(k_impro)) deallocate(k_impro, sign_impro, multiplicity_impro)
    allocate(k_impro(nit), sign_impro(nit), multiplicity_impro(nit))
    k_impro = k; sign_impro = sgn; multiplicity_impro = mlt
    if (allocated(ff_impro)) deallocate(ff_impro)
    allocate(ff_impro(np))
    do i = 1, np
       ff_impro(i) = coords(0.0, 0.0, 0.0)
    end do
    e_improper = 0.0d0
    v_improx = 0.0d0; v_improy = 0.0d0; v_improz = 0.0d0
    call impro_amber_force()
    call get_forces(ff_impro, fout)
    ei = e_improper
    vir(1) = v_improx; vir(2) = v_improy; vir(3) = v_improz
  end subroutine ref_improper_cvff

  ! dihed_PME_amber (Source/mod_force_dihed_amber.f90:7-252) with the 1-4 LJ tables zeroed,
  ! which leaves the torsion series alone in ff_dihed and e_dihedral.  The per-type
  ! tables are filled as the topology reader fills them (Source/mod_readfile.f90:405-410):
  ! k_dihed, multiplicity, cos_shift / sin_shift of the shift in degrees.
  subroutine ref_dihed_amber(n, lst, ntt, k, mult, shift_deg, fout, ed) &
       bind(c, name="ref_dihed_amber")
    integer(c_int), value :: n, ntt
    integer(c_int), intent(in) :: lst(5, n)
    real(c_double), intent(in) :: k(ntt), mult(ntt), shift_deg(ntt)
    real(c_float), intent(out) :: fout(3, np)
    real(c_double), intent(out) :: ed
    integer :: t
    if (allocated(dihedlist)) deallocate(dihedlist)
    allocate(dihedlist(5, n)); dihedlist = lst; numdihed = n
    if (allocated(k_dihed)) deallocate(k_dihed, shift_dihed, cos_shift, sin_shift, multiplicity)
    allocate(k_dihed(ntt), shift_dihed(ntt), cos_shift(ntt), sin_shift(ntt), multiplicity(ntt))
    do t = 1, ntt
       k_dihed(t) = k(t)
       shift_dihed(t) = shift_deg(t)
       cos_shift(t) = cos(acos(-1.0d0)*shift_deg(t)/180.0)
       sin_shift(t) = sin(acos(-1.0d0)*shift_deg(t)/180.0)
       multiplicity(t) = mult(t)
    end do
    lj14_1 = 0.0d0; lj14_2 = 0.0d0; lj14_3 = 0.0d0; lj14_4 = 0.0d0
    call zero_forces()
    e_dihedral = 0.0d0; pot_14 = 0.0d0
    v_dihedx = 0.0d0; v_dihedy = 0.0d0; v_dihedz = 0.0d0
    call dihed_PME_amber()
    call get_forces(ff_dihed, fout)
    ed = e_dihedral
  end subroutine ref_dihed_amber

  ! pbc (Source/mod_pbc.f90:6-39) on xyz in place.
  subroutine ref_pbc(xyz) bind(c, name="ref_pbc")
    real(c_float), intent(inout) :: xyz(3, np)
    integer :: i
    do i = 1, np
       position(i)%x = xyz(1, i); position(i)%y = xyz(2, i); position(i)%z = xyz(3, i)
    end do
    call pbc(0)
    do i = 1, np
       xyz(1, i) = position(i)%x; xyz(2, i) = position(i)%y; xyz(3, i) = position(i)%z
    end do
  end subroutine ref_pbc

end module sofia_ref_driver
