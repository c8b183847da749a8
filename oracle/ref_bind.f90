!---------------------------------------------------------------------!
! C-callable driver around the reference's own integral routines,     !
! compiled together with the reference sources (read in place from    !
! the reference tree by oracle/Makefile.ref) into oracle/_ref/.        !
!                                                                     !
! Test infrastructure only: it exists so that the tests can compare   !
! the C restatement (oracle/tcg_oracle.c) and the GPU engine with the !
! reference's own Fortran on the same inputs.  Nothing here is part   !
! of the engine.  Every routine sets the reference's module state     !
! (torus, unit cell, quadrature, keywords) explicitly, then calls the !
! reference routine unchanged:                                        !
!   ERI_shell_block_toroidal   src/Toroidal/ERI_shell_block_toroidal.f90:28
!   ERI_integral_toroidal      src/Toroidal/ERI_toroidal.f90:36        !
!   normalize_AO_toroidal_nD   src/basis/normalization/normalize_1D.f90:17
!   overlap/kinetic/nuclear    src/Toroidal/{1D,2D,3D}/*.f90           !
!   hartree_exchange_potential src/Energies/hartree_exchange_potential.f90:30
!   NRE                        src/Energies/Nuclear_repulsion.f90:1    !
!   fast_quad_self_check       src/Toroidal/fast_quad_self_check.f90:30 !
!   direct_init / direct_JK    src/module/eri_direct.f90:76,262        !
!   AO_to_MO_ERI_inplace, MP2  src/Tools/AO_to_MO/AO_to_MO.f90:20,     !
!                              src/Energies/MP2/MP2.f90:1              !
!   transform_ERI_AO_to_kMO,   src/Energies/SAO/MP2/AO_to_MO_SAO.f90:34 !
!   kMP2_calc                  src/Energies/SAO/MP2/MP2_SAO.f90:6      !
!   transform_AO_to_SAO,       src/Tools/SAO_Transfer.f90:1,108,235    !
!   transform_SAO_to_AO,                                               !
!   clean_and_symmetrize                                               !
!---------------------------------------------------------------------!
module ref_bind

      use iso_c_binding
      use classification_ERI
      use torus_init
      use unitcell_module
      use keywords
      use atom_basis
      use precomputed_bessel, only : nquad_point, fast_quad
      use clifford_tools,     only : init_clifford, setup_t_tables, pair_cache_free, &
                                     bessel_nmax_estimate, orbital_to_lxyz, t_bessel, &
                                     nq_t, t_node, t_weight, t_nmax, ps_on, ps_noscreen

      implicit none

      type(ERI_function), allocatable, save :: AOl(:)
      integer, save :: ndp_s = 1

      interface
        subroutine ERI_shell_block_toroidal(ndp,nA,SA,nB,SB,nC,SC,nD,SD,blk)
          import :: ERI_function
          integer, intent(in) :: ndp, nA, nB, nC, nD
          type(ERI_function), intent(in) :: SA(nA), SB(nB), SC(nC), SD(nD)
          double precision, intent(out) :: blk(nA,nB,nC,nD)
        end subroutine
        subroutine ERI_integral_toroidal(ndp,nBas,AO,two_electron_integrals)
          import :: ERI_function, ERI_size
          integer, intent(in) :: ndp, nBas
          type(ERI_function), intent(in) :: AO(nBas)
          double precision, intent(out) :: two_electron_integrals(ERI_size,nBas,nBas,nBas)
        end subroutine
      end interface

contains

      subroutine ref_init(ndp, Lv, ncl, nfuc_in, nq, fq, sao, phys) bind(C, name="ref_init")
        integer(c_int), value      :: ndp, nfuc_in, nq, fq, sao, phys
        real(c_double), intent(in) :: Lv(3)
        integer(c_int), intent(in) :: ncl(3)
        double precision, parameter :: pi = 3.14159265358979323846D00
        ndp_s = ndp
        Lx = Lv(1) ; Ly = Lv(2) ; Lz = Lv(3)
        ax = 2.d0*pi/Lx
        if (Ly > 0.d0) ay = 2.d0*pi/Ly
        if (Lz > 0.d0) az = 2.d0*pi/Lz
        nx = ncl(1) ; ny = ncl(2) ; nz = ncl(3)
        nfuc = nfuc_in
        N_cell = nx*ny*nz
        c_SAO = (sao /= 0)
        c_physicist = (phys /= 0)
        c_prim_screen = .true.
        c_symmetry = .false.
        c_debug = .false.
        c_fast_quad = (fq /= 0)
        fast_quad = (fq /= 0)
        nquad_point = nq
        ps_on = .false. ; ps_noscreen = .false.
        if (allocated(t_bessel)) deallocate(t_bessel)     ! force setup_t_tables to rebuild for this torus
        call pair_cache_free()
        call init_clifford()
      end subroutine ref_init

      subroutine label_of(l3, lab)
        integer, intent(in) :: l3(3)
        character(len=5), intent(out) :: lab
        character(len=1), parameter :: sh(0:4) = (/ 's','p','d','f','g' /)
        integer :: c, k
        lab = ' '
        lab(1:1) = sh(sum(l3))
        k = 1
        do c = 1, l3(1) ; k = k + 1 ; lab(k:k) = 'x' ; end do
        do c = 1, l3(2) ; k = k + 1 ; lab(k:k) = 'y' ; end do
        do c = 1, l3(3) ; k = k + 1 ; lab(k:k) = 'z' ; end do
      end subroutine label_of

      subroutine ref_set_basis(nbas, xyz, lxyz, nprim, expo, coef) bind(C, name="ref_set_basis")
        integer(c_int), value      :: nbas
        real(c_double), intent(in) :: xyz(3,nbas), expo(*), coef(*)
        integer(c_int), intent(in) :: lxyz(3,nbas), nprim(nbas)
        integer :: i, o
        if (allocated(AOl)) deallocate(AOl)
        allocate(AOl(nbas))
        o = 0
        do i = 1, nbas
          AOl(i)%x = xyz(1,i) ; AOl(i)%y = xyz(2,i) ; AOl(i)%z = xyz(3,i)
          call label_of(lxyz(:,i), AOl(i)%orbital)
          AOl(i)%exponent    = expo(o+1:o+nprim(i))
          AOl(i)%coefficient = coef(o+1:o+nprim(i))
          o = o + nprim(i)
        end do
      end subroutine ref_set_basis

      subroutine ref_get_coef(coef) bind(C, name="ref_get_coef")
        real(c_double), intent(out) :: coef(*)
        integer :: i, o, n
        o = 0
        do i = 1, size(AOl)
          n = size(AOl(i)%coefficient)
          coef(o+1:o+n) = AOl(i)%coefficient
          o = o + n
        end do
      end subroutine ref_get_coef

      subroutine ref_normalize() bind(C, name="ref_normalize")
        select case (ndp_s)
        case (1) ; call normalize_AO_toroidal_1D(size(AOl), AOl, .false.)
        case (2) ; call normalize_AO_toroidal_2D(size(AOl), AOl, .false.)
        case default ; call normalize_AO_toroidal_3D(size(AOl), AOl, .false.)
        end select
      end subroutine ref_normalize

      ! t-tables as the ERI driver builds them (src/Toroidal/ERI_toroidal.f90:94-104)
      subroutine ref_setup_tables(ncap_out) bind(C, name="ref_setup_tables")
        integer(c_int), intent(out) :: ncap_out
        integer :: i, lxyz(3), lmax, ncap
        double precision :: emax, amin, a_dir(3)
        lmax = 0 ; emax = 0.d0
        do i = 1, size(AOl)
          call orbital_to_lxyz(AOl(i)%orbital, lxyz)
          lmax = max(lmax, maxval(lxyz(1:ndp_s)))
          emax = max(emax, maxval(AOl(i)%exponent))
        end do
        a_dir = (/ ax, ay, az /)
        amin = minval(a_dir(1:ndp_s))
        ncap = bessel_nmax_estimate(4.d0*emax/(amin*amin)) + 2*lmax + 1
        call setup_t_tables(ncap, ndp_s)
        ncap_out = ncap
      end subroutine ref_setup_tables

      subroutine ref_get_nodes(nq, t, w, tnmax) bind(C, name="ref_get_nodes")
        integer(c_int), intent(out) :: nq
        real(c_double), intent(out) :: t(*), w(*)
        integer(c_int), intent(out) :: tnmax(*)
        integer :: k, d
        nq = nq_t
        t(1:nq_t) = t_node(1:nq_t)
        w(1:nq_t) = t_weight(1:nq_t)
        do d = 1, ndp_s
          do k = 1, nq_t
            tnmax(k + nq_t*(d-1)) = t_nmax(k,d)
          end do
        end do
      end subroutine ref_get_nodes

      ! one shell block; ia..id are 1-based first AOs, na..nd component counts
      subroutine ref_eri_block(ia, na, ib, nb, ic, nc, id, nd, blk) bind(C, name="ref_eri_block")
        integer(c_int), value       :: ia, na, ib, nb, ic, nc, id, nd
        real(c_double), intent(out) :: blk(na,nb,nc,nd)
        call ERI_shell_block_toroidal(ndp_s, na, AOl(ia:ia+na-1), nb, AOl(ib:ib+nb-1), &
                                      nc, AOl(ic:ic+nc-1), nd, AOl(id:id+nd-1), blk)
      end subroutine ref_eri_block

      subroutine ref_eri_full(eri) bind(C, name="ref_eri_full")
        real(c_double), intent(out) :: eri(*)
        integer :: nb
        nb = size(AOl)
        if (c_SAO) then
          ERI_size = nfuc
        else
          ERI_size = nb
        end if
        call ERI_integral_toroidal(ndp_s, nb, AOl, eri)
      end subroutine ref_eri_full

      subroutine ref_one_electron(natom, geom, zat, S, T, V) bind(C, name="ref_one_electron")
        integer(c_int), value       :: natom
        real(c_double), intent(in)  :: geom(natom,3)
        integer(c_int), intent(in)  :: zat(natom)
        real(c_double), intent(out) :: S(*), T(*), V(*)
        type(atom), allocatable :: ats(:)
        integer :: i, nb
        nb = size(AOl)
        allocate(ats(natom))
        do i = 1, natom
          ats(i)%charge = zat(i)
        end do
        select case (ndp_s)
        case (1)
          call overlap_matrix_toroidal_1D(nb, AOl, S)
          call kinetic_matrix_toroidal_1D(nb, AOl, T)
          call nuclear_attraction_matrix_toroidal_1D(natom, nb, geom, ats, AOl, V)
        case (2)
          call overlap_matrix_toroidal_2D(nb, AOl, S)
          call kinetic_matrix_toroidal_2D(nb, AOl, T)
          call nuclear_attraction_matrix_toroidal_2D(natom, nb, geom, ats, AOl, V)
        case default
          call overlap_matrix_toroidal_3D(nb, AOl, S)
          call kinetic_matrix_toroidal_3D(nb, AOl, T)
          call nuclear_attraction_matrix_toroidal_3D(natom, nb, geom, ats, AOl, V)
        end select
      end subroutine ref_one_electron

      ! The keyword Details: details_integrals (full ERI(nBas,nBas,nBas,nBas)) or
      ! details_integrals_SAO (ERI(nfuc,nBas,nBas,nBas)) write OV.dat, KI.dat,
      ! NA.dat, the *_matrix.dat files and ERI.dat into the directory 'dir'
      ! (module variable tmp_file_name)
      subroutine ref_details(nbas, sao, S, T, V, eri, dir) bind(C, name="ref_details")
        use files, only : tmp_file_name
        integer(c_int), value      :: nbas, sao
        real(c_double), intent(in) :: S(nbas,nbas), T(nbas,nbas), V(nbas,nbas), eri(*)
        character(kind=c_char), intent(in) :: dir(*)
        integer :: i
        tmp_file_name = ' '
        i = 1
        do while (dir(i) /= c_null_char .and. i <= 100)
          tmp_file_name(i:i) = dir(i)
          i = i + 1
        end do
        if (sao /= 0) then
          call details_integrals_SAO(nbas, S, T, V, eri)
        else
          call details_integrals(nbas, S, T, V, eri)
        end if
      end subroutine ref_details

      ! Gamma MP2: AO_to_MO_ERI_inplace on eri (overwritten), then MP2, whose
      ! report goes to the file 'path' (unit outfile)
      subroutine ref_mp2(nbas, nocc, e, cmo, eri, path) bind(C, name="ref_mp2")
        use files, only : outfile
        integer(c_int), value         :: nbas, nocc
        real(c_double), intent(in)    :: e(nbas), cmo(nbas,nbas)
        real(c_double), intent(inout) :: eri(nbas,nbas,nbas,nbas)
        character(kind=c_char), intent(in) :: path(*)
        character(len=512) :: fname
        integer :: i
        fname = ' '
        i = 1
        do while (path(i) /= c_null_char .and. i <= 512)
          fname(i:i) = path(i)
          i = i + 1
        end do
        open(unit=outfile, file=trim(fname), status='replace', action='write')
        call AO_to_MO_ERI_inplace(nbas, cmo, eri, .false.)
        call MP2(nbas, nocc, e, eri, 0.d0, 0.d0, .false.)
        close(outfile)
      end subroutine ref_mp2

      ! k-MP2 from the SAO ERI(nfuc,nBas,nBas,nBas): k blocks built from the
      ! block-diagonal coefficients, energies and occupations (2 or 0) given per
      ! k (ordered as the cells), the (ov|ov) ranges as set_cc_dims sets them
      ! (src/module/cc_dims.f90:68-88), transform_ERI_AO_to_kMO, kMP2_calc.
      ! Its report goes to the file 'path' (unit outfile).
      subroutine ref_kmp2(nbas, csao, ek, occ, eri, path, ec) bind(C, name="ref_kmp2")
        use files, only : outfile
        use k_blocks, only : K_block
        integer(c_int), value          :: nbas
        complex(c_double_complex), intent(in) :: csao(nbas,nbas)
        real(c_double), intent(in)     :: ek(nbas), eri(nfuc,nbas,nbas,nbas)
        integer(c_int), intent(in)     :: occ(nbas)
        character(kind=c_char), intent(in) :: path(*)
        real(c_double), intent(out)    :: ec
        type(K_block), allocatable :: blk(:)
        complex(8), allocatable :: ekmo(:,:,:,:,:,:,:)
        character(len=512) :: fname
        integer :: i, k, o_hi, omin, nok
        interface
          subroutine transform_ERI_AO_to_kMO(nBas, ERI_AO, coeff_SAO, p_lo, p_hi, q_lo, q_hi, ERI_out)
            use unitcell_module, only : nfuc, N_cell
            integer, intent(in) :: nBas, p_lo, p_hi, q_lo, q_hi
            real(8), intent(in) :: ERI_AO(nfuc,nBas,nBas,nBas)
            complex(8), intent(in) :: coeff_SAO(nBas,nBas)
            complex(8), intent(out) :: ERI_out(p_lo:p_hi,q_lo:q_hi,p_lo:p_hi,q_lo:q_hi,N_cell,N_cell,N_cell)
          end subroutine
          subroutine kMP2_calc(N_cell, blk, p_lo, p_hi, q_lo, q_hi, ERI_kMO, ENuc, EHF, EcMP2)
            use k_blocks, only : K_block
            integer, intent(in) :: N_cell, p_lo, p_hi, q_lo, q_hi
            type(K_block), intent(in) :: blk(N_cell)
            complex(8), intent(in) :: ERI_kMO(p_lo:p_hi,q_lo:q_hi,p_lo:p_hi,q_lo:q_hi,N_cell,N_cell,N_cell)
            real(8), intent(in) :: ENuc, EHF
            real(8), intent(out) :: EcMP2
          end subroutine
        end interface
        fname = ' '
        i = 1
        do while (path(i) /= c_null_char .and. i <= 512)
          fname(i:i) = path(i)
          i = i + 1
        end do
        allocate(blk(N_cell))
        o_hi = 0 ; omin = nfuc
        do k = 1, N_cell
          blk(k)%dim = nfuc
          allocate(blk(k)%occupation(nfuc), blk(k)%energy(nfuc), blk(k)%Coeff(nfuc,nfuc))
          blk(k)%occupation = occ((k-1)*nfuc+1:k*nfuc)
          blk(k)%energy = ek((k-1)*nfuc+1:k*nfuc)
          blk(k)%Coeff = csao((k-1)*nfuc+1:k*nfuc, (k-1)*nfuc+1:k*nfuc)
          nok = count(blk(k)%occupation == 2)
          o_hi = max(o_hi, nok) ; omin = min(omin, nok)
        end do
        allocate(ekmo(1:o_hi, omin+1:nfuc, 1:o_hi, omin+1:nfuc, N_cell, N_cell, N_cell))
        open(unit=outfile, file=trim(fname), status='replace', action='write')
        call transform_ERI_AO_to_kMO(nbas, eri, csao, 1, o_hi, omin+1, nfuc, ekmo)
        call kMP2_calc(N_cell, blk, 1, o_hi, omin+1, nfuc, ekmo, 0.d0, 0.d0, ec)
        close(outfile)
      end subroutine ref_kmp2

      ! transform_AO_to_SAO(nBas, AO, SAO) on the unit cell set by ref_init
      subroutine ref_ao_to_sao(nbas, ao, sao) bind(C, name="ref_ao_to_sao")
        integer(c_int), value          :: nbas
        real(c_double), intent(in)     :: ao(nbas,nbas)
        complex(c_double_complex), intent(out) :: sao(nbas,nbas)
        interface
          subroutine transform_AO_to_SAO(nBas, AO_matrix, SAO_matrix)
            integer, intent(in) :: nBas
            real(8), intent(in) :: AO_matrix(nBas,nBas)
            complex(8), intent(out) :: SAO_matrix(nBas,nBas)
          end subroutine
        end interface
        call transform_AO_to_SAO(nbas, ao, sao)
      end subroutine ref_ao_to_sao

      ! transform_SAO_to_AO(nBas, SAO, AO)
      subroutine ref_sao_to_ao(nbas, sao, ao) bind(C, name="ref_sao_to_ao")
        integer(c_int), value          :: nbas
        complex(c_double_complex), intent(in) :: sao(nbas,nbas)
        real(c_double), intent(out)    :: ao(nbas,nbas)
        interface
          subroutine transform_SAO_to_AO(nBas, M_SAO, M_AO)
            integer, intent(in) :: nBas
            complex(8), intent(in) :: M_SAO(nBas,nBas)
            real(8), intent(out) :: M_AO(nBas,nBas)
          end subroutine
        end interface
        call transform_SAO_to_AO(nbas, sao, ao)
      end subroutine ref_sao_to_ao

      ! clean_and_symmetrize(N, A) in place
      subroutine ref_clean_and_symmetrize(n, a) bind(C, name="ref_clean_and_symmetrize")
        integer(c_int), value          :: n
        complex(c_double_complex), intent(inout) :: a(n,n)
        interface
          subroutine clean_and_symmetrize(N, A)
            integer, intent(in) :: N
            complex(8), intent(inout) :: A(N,N)
          end subroutine
        end interface
        call clean_and_symmetrize(n, a)
      end subroutine ref_clean_and_symmetrize

      ! direct J/K on the density P(nBas,nBas): direct_init then direct_JK
      subroutine ref_direct_jk(P, J, K) bind(C, name="ref_direct_jk")
        use eri_direct, only : direct_init, direct_JK
        real(c_double), intent(in)  :: P(size(AOl),size(AOl))
        real(c_double), intent(out) :: J(size(AOl),size(AOl)), K(size(AOl),size(AOl))
        call direct_init(ndp_s, size(AOl), AOl)
        call direct_JK(size(AOl), P, J, K)
      end subroutine ref_direct_jk

      ! the start-up self-check; its report line goes to the file 'path'
      ! (unit outfile); out2 = (fast_quad after the check, nquad_point after)
      subroutine ref_fast_quad_self_check(natom, geom, zat, nq128, path, out2) bind(C, name="ref_fast_quad_self_check")
        use files, only : outfile
        integer(c_int), value       :: natom, nq128
        real(c_double), intent(in)  :: geom(natom,3)
        integer(c_int), intent(in)  :: zat(natom)
        character(kind=c_char), intent(in) :: path(*)
        real(c_double), intent(out) :: out2(2)
        type(atom), allocatable :: ats(:)
        character(len=512) :: fname
        integer :: i
        fname = ' '
        i = 1
        do while (path(i) /= c_null_char .and. i <= 512)
          fname(i:i) = path(i)
          i = i + 1
        end do
        allocate(ats(natom))
        do i = 1, natom
          ats(i)%charge = zat(i)
        end do
        c_nquad_128 = (nq128 /= 0)
        open(unit=outfile, file=trim(fname), status='replace', action='write')
        call fast_quad_self_check(ndp_s, size(AOl), AOl, natom, geom, ats)
        close(outfile)
        out2(1) = merge(1.d0, 0.d0, fast_quad)
        out2(2) = dble(nquad_point)
      end subroutine ref_fast_quad_self_check

      subroutine ref_nre(natom, geom, zat, E) bind(C, name="ref_nre")
        integer(c_int), value       :: natom
        real(c_double), intent(in)  :: geom(natom,3)
        integer(c_int), intent(in)  :: zat(natom)
        real(c_double), intent(out) :: E
        type(atom), allocatable :: ats(:)
        character(len=10) :: ctype
        integer :: i
        allocate(ats(natom))
        do i = 1, natom
          ats(i)%charge = zat(i)
        end do
        select case (ndp_s)
        case (1) ; ctype = 'Tori1D'
        case (2) ; ctype = 'Tori2D'
        case default ; ctype = 'Tori3D'
        end select
        call NRE(ctype, natom, geom, ats, E)
      end subroutine ref_nre

      ! scaled Bessel function and vectors (src/module/bessel_iv_log.f90:320,
      ! src/module/clifford_tools.f90:386,429), quadrature nodes (precompute_t.f90:906)
      real(c_double) function ref_iv_scaled(v, x) bind(C, name="ref_iv_scaled")
        use bessel_functions, only : iv_scaled
        integer(c_int), value :: v
        real(c_double), value :: x
        ref_iv_scaled = iv_scaled(v, x)
      end function ref_iv_scaled

      subroutine ref_bessel_vector(n, A, fast, out) bind(C, name="ref_bessel_vector")
        use clifford_tools, only : bessel_scaled_vector, bessel_scaled_vector_fast
        integer(c_int), value :: n, fast
        real(c_double), value :: A
        real(c_double), intent(out) :: out(0:n)
        if (fast /= 0) then
          call bessel_scaled_vector_fast(n, A, out)
        else
          call bessel_scaled_vector(n, A, out)
        end if
      end subroutine ref_bessel_vector

      subroutine ref_fast_quad_points(n, t, w) bind(C, name="ref_fast_quad_points")
        use precomputed_bessel, only : fast_quad_points
        integer(c_int), value :: n
        real(c_double), intent(out) :: t(n), w(n)
        call fast_quad_points(n, t, w)
      end subroutine ref_fast_quad_points

      subroutine ref_bary(d, e1, e2, x1, x2, pe, cen) bind(C, name="ref_bary")
        use clifford_tools, only : bary_dir
        integer(c_int), value :: d
        real(c_double), value :: e1, e2, x1, x2
        real(c_double), intent(out) :: pe, cen
        call bary_dir(d, e1, e2, x1, x2, pe, cen)
      end subroutine ref_bary

      subroutine ref_jk(n1, nbas, P, eri, J, K) bind(C, name="ref_jk")
        integer(c_int), value       :: n1, nbas
        real(c_double), intent(in)  :: P(nbas,nbas), eri(*)
        real(c_double), intent(out) :: J(nbas,nbas), K(nbas,nbas)
        call hartree_exchange_potential(n1, nbas, P, eri, J, K)
      end subroutine ref_jk

end module ref_bind
