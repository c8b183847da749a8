! tcg_engine.f90 -- ISO_C_BINDING interfaces of the MI355X Clifford-Gaussian
! engine (include/tcg.h) for the reference's Fortran driver.  Compile it with
! the driver's other modules (src/module/) and link with
!   -L<repo>/william_ci_amd -ltcg -Wl,-rpath,<repo>/william_ci_amd
!
! Arrays are the caller's, column-major FP64, exactly as the reference
! declares them; every function returns 0 on success (TCG_OK) or a negative
! status with a message from tcg_message(tcg_last_error(ctx)).  Arguments that
! may be NULL on the C side (the ERI of tcg_eri_build and tcg_jk_stored*) are
! type(c_ptr), value: pass c_loc(ERI) for a host copy, c_null_ptr to keep the
! tensor on the GPU only (or to use the resident one).
module tcg_engine
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TCG_OK = 0, TCG_E_ARG = -1, TCG_E_STATE = -2, TCG_E_DEVICE = -3, &
                               TCG_E_LATTICE = -4, TCG_E_NOMEM = -5

  type(c_ptr), save :: tcg = c_null_ptr        ! one context per torus system
  type(c_ptr), save :: tcg_dfc = c_null_ptr    ! density fit (keyword density_fitting)

  interface
    ! torus, cells and fast rule: Torus_init / unitcell_module / nquad_point
    integer(c_int) function tcg_create(ndp, L, ncell, nquad, ctx) bind(C, name="tcg_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ndp, nquad
      real(c_double), intent(in) :: L(3)
      integer(c_int), intent(in) :: ncell(3)
      type(c_ptr), intent(out) :: ctx
    end function
    ! AO list as structure of arrays (type(ERI_function) AO(nBas), normalised coefficients)
    integer(c_int) function tcg_set_basis(ctx, nbas, nfuc, center, lxyz, prim_off, expo, coef) &
        bind(C, name="tcg_set_basis")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas, nfuc
      real(c_double), intent(in) :: center(3,*), expo(*), coef(*)
      integer(c_int), intent(in) :: lxyz(3,*), prim_off(*)
    end function
    ! ERI_integral_toroidal(ndp,nBas,AO,ERI); eri = c_loc(ERI) or c_null_ptr (GPU only)
    integer(c_int) function tcg_eri_build(ctx, sao, physicist, eri) bind(C, name="tcg_eri_build")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, eri
      integer(c_int), value :: sao, physicist
    end function
    ! hartree_exchange_potential(n1,nBas,P,ERI,J,K); eri = c_null_ptr: the resident tensor
    integer(c_int) function tcg_jk_stored(ctx, n1, nbas, P, eri, J, K) bind(C, name="tcg_jk_stored")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, eri
      integer(c_int), value :: n1, nbas
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function
    integer(c_int) function tcg_jk_stored_order(ctx, n1, nbas, P, eri, physicist, J, K) &
        bind(C, name="tcg_jk_stored_order")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, eri
      integer(c_int), value :: n1, nbas, physicist
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function
    ! the same from the resident tensor with no device allocation (callers that cannot allocate)
    integer(c_int) function tcg_jk_stored_noalloc(ctx, n1, nbas, P, J, K) bind(C, name="tcg_jk_stored_noalloc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1, nbas
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function
    ! direct_init / direct_JK
    integer(c_int) function tcg_direct_init(ctx) bind(C, name="tcg_direct_init")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tcg_direct_jk(ctx, nbas, P, J, K) bind(C, name="tcg_direct_jk")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function
    ! AO_to_MO_ERI_inplace + MP2 (Gamma supercell)
    integer(c_int) function tcg_mp2_gamma(ctx, nbas, nocc, e, C, out) bind(C, name="tcg_mp2_gamma")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas, nocc
      real(c_double), intent(in) :: e(*), C(*)
      real(c_double), intent(out) :: out(4)
    end function
    ! transform_ERI_AO_to_kMO + kMP2_calc
    integer(c_int) function tcg_kmp2(ctx, nbas, coeff_sao, e, occ, out) bind(C, name="tcg_kmp2")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      complex(c_double_complex), intent(in) :: coeff_sao(*)
      real(c_double), intent(in) :: e(*)
      integer(c_int), intent(in) :: occ(*)
      real(c_double), intent(out) :: out(5)
    end function
    ! transform_ERI_direct_kMO + kMP2_calc (integral-direct)
    integer(c_int) function tcg_kmp2_direct(ctx, nbas, coeff_sao, e, occ, slab_cells, out) &
        bind(C, name="tcg_kmp2_direct")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas, slab_cells
      complex(c_double_complex), intent(in) :: coeff_sao(*)
      real(c_double), intent(in) :: e(*)
      integer(c_int), intent(in) :: occ(*)
      real(c_double), intent(out) :: out(8)
    end function
    ! transform_AO_to_SAO / transform_SAO_to_AO
    integer(c_int) function tcg_ao_to_sao(ctx, nbas, M_ao, M_sao) bind(C, name="tcg_ao_to_sao")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      real(c_double), intent(in) :: M_ao(*)
      complex(c_double_complex), intent(out) :: M_sao(*)
    end function
    integer(c_int) function tcg_sao_to_ao(ctx, nbas, M_sao, M_ao) bind(C, name="tcg_sao_to_ao")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      complex(c_double_complex), intent(in) :: M_sao(*)
      real(c_double), intent(out) :: M_ao(*)
    end function
    ! diagonalize_block_diagonal_complex_para
    integer(c_int) function tcg_heev_blocks(ctx, nbas, A_sao, e) bind(C, name="tcg_heev_blocks")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      complex(c_double_complex), intent(inout) :: A_sao(*)
      real(c_double), intent(out) :: e(*)
    end function
    ! RHF_SAO (the whole iteration on the GPU)
    integer(c_int) function tcg_rhf_sao(ctx, source, nbas, nocc, S, T, V, max_scf, thresh, diis, e, &
        coeff_sao, occ, P, out) bind(C, name="tcg_rhf_sao")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: source, nbas, nocc, max_scf, diis
      real(c_double), value :: thresh
      real(c_double), intent(in) :: S(*), T(*), V(*)
      real(c_double), intent(out) :: e(*), P(*), out(12)
      complex(c_double_complex), intent(out) :: coeff_sao(*)
      integer(c_int), intent(out) :: occ(*)
    end function
    ! df_init, df_JK, df_kMO + kMP2_calc
    integer(c_int) function tcg_df_create(ctx, nau_cell, aux_center, aux_lxyz, aux_prim_off, aux_expo, &
        aux_coef, df) bind(C, name="tcg_df_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nau_cell
      real(c_double), intent(in) :: aux_center(3,*), aux_expo(*), aux_coef(*)
      integer(c_int), intent(in) :: aux_lxyz(3,*), aux_prim_off(*)
      type(c_ptr), intent(out) :: df
    end function
    integer(c_int) function tcg_df_info(df, info, dmin) bind(C, name="tcg_df_info")
      import :: c_int, c_double, c_long_long, c_ptr
      type(c_ptr), value :: df
      integer(c_long_long), intent(out) :: info(3)
      real(c_double), intent(out) :: dmin
    end function
    integer(c_int) function tcg_df_b(df, B) bind(C, name="tcg_df_b")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: df
      real(c_double), intent(out) :: B(*)
    end function
    integer(c_int) function tcg_df_jk(df, nbas, P, J, K) bind(C, name="tcg_df_jk")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: df
      integer(c_int), value :: nbas
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function
    integer(c_int) function tcg_df_kmp2(df, nbas, coeff_sao, e, occ, out) bind(C, name="tcg_df_kmp2")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: df
      integer(c_int), value :: nbas
      complex(c_double_complex), intent(in) :: coeff_sao(*)
      real(c_double), intent(in) :: e(*)
      integer(c_int), intent(in) :: occ(*)
      real(c_double), intent(out) :: out(5)
    end function
    integer(c_int) function tcg_df_kmp2_fp32(df, nbas, coeff_sao, e, occ, out) bind(C, name="tcg_df_kmp2_fp32")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: df
      integer(c_int), value :: nbas
      complex(c_double_complex), intent(in) :: coeff_sao(*)
      real(c_double), intent(in) :: e(*)
      integer(c_int), intent(in) :: occ(*)
      real(c_double), intent(out) :: out(5)
    end function
    integer(c_int) function tcg_df_times(df, out) bind(C, name="tcg_df_times")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: df
      real(c_double), intent(out) :: out(3)
    end function
    type(c_ptr) function tcg_df_last_error(df) bind(C, name="tcg_df_last_error")
      import :: c_ptr
      type(c_ptr), value :: df
    end function
    integer(c_int) function tcg_df_destroy(df) bind(C, name="tcg_df_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: df
    end function
    ! overlap / kinetic / nuclear_attraction_matrix_toroidal_nD
    integer(c_int) function tcg_one_electron(ctx, natom, xyz, Z, S, T, V) bind(C, name="tcg_one_electron")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: natom
      real(c_double), intent(in) :: xyz(3,*)
      integer(c_int), intent(in) :: Z(*)
      real(c_double), intent(out) :: S(*), T(*), V(*)
    end function
    ! fast_quad / nquad_point and the start-up self-check
    ! normalize_AO_toroidal_nD
    integer(c_int) function tcg_normalize(ctx, nbas, center, lxyz, prim_off, expo, coef_raw, coef_out) &
        bind(C, name="tcg_normalize")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      real(c_double), intent(in) :: center(3,*), expo(*), coef_raw(*)
      integer(c_int), intent(in) :: lxyz(3,*), prim_off(*)
      real(c_double), intent(out) :: coef_out(*)
    end function
    integer(c_int) function tcg_set_quadrature(ctx, fast, nquad) bind(C, name="tcg_set_quadrature")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: fast, nquad
    end function
    integer(c_int) function tcg_get_quadrature(ctx, fast, nquad, t, w) bind(C, name="tcg_get_quadrature")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, w
      integer(c_int), intent(out) :: fast, nquad
    end function
    integer(c_int) function tcg_fast_quad_self_check(ctx, natom, xyz, Z, nq_full, deri, dv, fast_on) &
        bind(C, name="tcg_fast_quad_self_check")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: natom, nq_full
      real(c_double), intent(in) :: xyz(3,*)
      integer(c_int), intent(in) :: Z(*)
      real(c_double), intent(out) :: deri, dv
      integer(c_int), intent(out) :: fast_on
    end function
    ! ERI_shell_block_toroidal over a list of shell quartets
    integer(c_int) function tcg_eri_blocks(ctx, n, shells, out, offsets) bind(C, name="tcg_eri_blocks")
      import :: c_int, c_double, c_long_long, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int), intent(in) :: shells(4,*)
      real(c_double), intent(out) :: out(*)
      integer(c_long_long), intent(out) :: offsets(*)
    end function
    ! learned per-quartet orders (no reference counterpart)
    integer(c_int) function tcg_set_learned(ctx, params, norm, orders, norders, margin) &
        bind(C, name="tcg_set_learned")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: params(*), norm(24)
      integer(c_int), intent(in) :: orders(*)
      integer(c_int), value :: norders
      real(c_float), value :: margin
    end function
    integer(c_int) function tcg_set_learned_skip(ctx, thresh, schwarz_cap, margin_b) &
        bind(C, name="tcg_set_learned_skip")
      import :: c_int, c_double, c_float, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: thresh, schwarz_cap
      real(c_float), value :: margin_b
    end function
    integer(c_int) function tcg_eri_build_learned(ctx, sao, physicist, eri, rule_counts) &
        bind(C, name="tcg_eri_build_learned")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, eri, rule_counts
      integer(c_int), value :: sao, physicist
    end function
    integer(c_int) function tcg_direct_init_learned(ctx, rule_counts) bind(C, name="tcg_direct_init_learned")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rule_counts
    end function
    integer(c_int) function tcg_set_direct_screen(ctx, tau, full_every, use_learned_bound) &
        bind(C, name="tcg_set_direct_screen")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: tau
      integer(c_int), value :: full_every, use_learned_bound
    end function
    integer(c_int) function tcg_direct_jk_screened(ctx, nbas, P, tau, J, K) bind(C, name="tcg_direct_jk_screened")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbas
      real(c_double), value :: tau
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function
    integer(c_int) function tcg_direct_screen_info(ctx, tau, full_every, counts) bind(C, name="tcg_direct_screen_info")
      import :: c_int, c_double, c_long_long, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tau
      integer(c_int), intent(out) :: full_every
      integer(c_long_long), intent(out) :: counts(5)
    end function
    integer(c_int) function tcg_mlp_param_count() bind(C, name="tcg_mlp_param_count")
      import :: c_int
    end function
    integer(c_int) function tcg_mlp_forward(params, X, n, Y) bind(C, name="tcg_mlp_forward")
      import :: c_int, c_float
      real(c_float), intent(in) :: params(*), X(*)
      integer(c_int), value :: n
      real(c_float), intent(out) :: Y(*)
    end function
    integer(c_int) function tcg_mlp_grad(params, X, T, n, w, grad, loss) bind(C, name="tcg_mlp_grad")
      import :: c_int, c_float
      real(c_float), intent(in) :: params(*), X(*), T(*), w(2)
      integer(c_int), value :: n
      real(c_float), intent(out) :: grad(*), loss
    end function
    integer(c_int) function tcg_mlp_train(params, X, T, n, epochs, batch, lr, seed, w, loss_hist) &
        bind(C, name="tcg_mlp_train")
      import :: c_int, c_float, c_ptr
      real(c_float), intent(inout) :: params(*)
      real(c_float), intent(in) :: X(*), T(*), w(2)
      integer(c_int), value :: n, epochs, batch, seed
      real(c_float), value :: lr
      type(c_ptr), value :: loss_hist
    end function
    ! counters, timings and the resident tensor
    integer(c_int) function tcg_eri_stats(ctx, out, n) bind(C, name="tcg_eri_stats")
      import :: c_int, c_long_long, c_ptr
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: out(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function tcg_eri_times(ctx, out, n) bind(C, name="tcg_eri_times")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function tcg_eri_device(ctx, dptr, count) bind(C, name="tcg_eri_device")
      import :: c_int, c_long_long, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: dptr
      integer(c_long_long), intent(out) :: count
    end function
    type(c_ptr) function tcg_last_error(ctx) bind(C, name="tcg_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tcg_destroy(ctx) bind(C, name="tcg_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_size_t) function tcg_strlen_(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! The C string a tcg_last_error / tcg_df_last_error call returned, as Fortran text
  ! (for the driver's "error stop" on a non-zero status).
  function tcg_message(p) result(msg)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(tcg_strlen_(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function tcg_message

end module tcg_engine
