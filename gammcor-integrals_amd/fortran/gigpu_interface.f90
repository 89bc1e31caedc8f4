! ISO_C_BINDING interface to libgigpu (include/gigpu.h) for the Fortran host.
!
! Module gigpu_c declares every C entry point.  Module gigpu_cholesky wraps the
! two calls Gammcor makes, chol_gammcor_Rkpq and chol_gammcor_Rkab
! (Cholesky_Gammcor.f90:26-92,335-377), over plain arrays: the TAOBasis
! components (basis_definitions.f90:94-192) go in unchanged, Rkpq and the
! TChol2Vecs index arrays come back allocated by the wrapper, as the
! reference allocates them inside the callee.  Any failure of the library
! ends in error stop with the library's message, which is what the reference
! does through msg(..., MSG_ERROR).  Integers are 64-bit throughout (the
! reference is built with -i8).  Fortran 2008, no coarrays.
module gigpu_c
  use iso_c_binding
  implicit none

  integer(c_int64_t), parameter :: GI_CHOL_ACCURACY_DEFAULT = 1
  integer(c_int64_t), parameter :: GI_CHOL_ACCURACY_TIGHT = 2
  integer(c_int64_t), parameter :: GI_CHOL_ACCURACY_LUDICROUS = 3
  integer(c_int64_t), parameter :: GI_CHOL_ACCURACY_DEBUG = 4
  integer(c_int64_t), parameter :: GI_CHOL_DECODE_FIXED = 1
  integer(c_int64_t), parameter :: GI_CHOL_SKIP_VECTORS = 2
  integer(c_int64_t), parameter :: GI_CHOL_TIME_ERI = 4
  integer(c_int64_t), parameter :: GI_CHOL_RAW_TIES = 8
  integer(c_int64_t), parameter :: GI_MO_ALL_COLUMNS = 1

  interface
    integer(c_int) function gi_create(ctx, device) bind(C, name="gi_create")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), intent(out) :: ctx
      integer(c_int64_t), value :: device
    end function
    integer(c_int) function gi_destroy(ctx) bind(C, name="gi_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function gi_last_error() bind(C, name="gi_last_error")
      import :: c_ptr
    end function
    integer(c_int) function gi_set_basis(ctx, natoms, coords, nshells, shell_centers, shell_params, nparams, &
        momentum, nprim, max_nprim, cntr, exps, max_nang, norm, spher) bind(C, name="gi_set_basis")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: natoms, nshells, nparams, max_nprim, max_nang, spher
      real(c_double), intent(in) :: coords(3, *), cntr(max_nprim, *), exps(max_nprim, *), norm(max_nang, *)
      integer(c_int64_t), intent(in) :: shell_centers(*), shell_params(*), momentum(*), nprim(*)
    end function
    integer(c_int) function gi_eri_quartets(ctx, nq, quartets, kappa, offsets, out, out_len) &
        bind(C, name="gi_eri_quartets")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nq, out_len
      integer(c_int64_t), intent(in) :: quartets(4, *), offsets(*)
      real(c_double), value :: kappa
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function gi_eri_pair_block(ctx, nrow_pairs, row_pairs, ncol_pairs, col_pairs, kappa, out, ldo) &
        bind(C, name="gi_eri_pair_block")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nrow_pairs, ncol_pairs, ldo
      integer(c_int64_t), intent(in) :: row_pairs(2, *), col_pairs(2, *)
      real(c_double), value :: kappa
      real(c_double), intent(out) :: out(ldo, *)
    end function
    integer(c_int) function gi_set_option(ctx, name, value) bind(C, name="gi_set_option")
      import :: c_int, c_ptr, c_int64_t, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t), value :: value
    end function
    integer(c_int) function gi_chol_build(ctx, tau, kappa, max_block_dim, vec_block, flags) &
        bind(C, name="gi_chol_build")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tau, kappa
      integer(c_int64_t), value :: max_block_dim, vec_block, flags
    end function
    integer(c_int) function gi_chol_build_batch(n, ctxs, tau, kappa, max_block_dim, vec_block, flags, nworkers, &
        status) bind(C, name="gi_chol_build_batch")
      import :: c_int, c_ptr, c_int64_t, c_double
      integer(c_int64_t), value :: n, max_block_dim, vec_block, flags, nworkers
      type(c_ptr), intent(in) :: ctxs(*)
      real(c_double), value :: tau, kappa
      integer(c_int64_t), intent(out) :: status(*)
    end function
    integer(c_int) function gi_one_electron(ctx, atom_charges, npc, pc_coords, pc_charges, S, T, V, ldh) &
        bind(C, name="gi_one_electron")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, atom_charges, pc_coords, pc_charges, S, T, V
      integer(c_int64_t), value :: npc, ldh
    end function
    integer(c_int) function gi_h0(ctx, atom_charges, npc, pc_coords, pc_charges, H0, ldh) &
        bind(C, name="gi_h0")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, pc_coords, pc_charges
      real(c_double), intent(in) :: atom_charges(*)
      integer(c_int64_t), value :: npc, ldh
      real(c_double), intent(out) :: H0(ldh, *)
    end function
    integer(c_int) function gi_chol_residual(ctx, max_residual, min_residual) bind(C, name="gi_chol_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: max_residual, min_residual
    end function
    integer(c_int) function gi_coulomb_energy_grad(ctx, density, ldd, kappa, energy, grad) &
        bind(C, name="gi_coulomb_energy_grad")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ldd
      real(c_double), intent(in) :: density(ldd, *)
      real(c_double), value :: kappa
      real(c_double), intent(out) :: energy
      type(c_ptr), value :: grad
    end function
    integer(c_int) function gi_chol_accuracy_tau(accuracy, tau) bind(C, name="gi_chol_accuracy_tau")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: accuracy
      real(c_double), intent(out) :: tau
    end function
    integer(c_int) function gi_chol_dims(ctx, dims) bind(C, name="gi_chol_dims")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: dims(8)
    end function
    integer(c_int) function gi_chol_meta(ctx, shell_pairs, shell_pair_loc, shell_pair_dim, subset_dim, &
        subset_bounds, pivots, pivot_order, pivot_shell_pairs, pivot_shell_pair_loc, pivot_shell_pair_dim, &
        pivot_orb_pairs) bind(C, name="gi_chol_meta")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: shell_pairs(2, *), shell_pair_loc(3, *), shell_pair_dim(*), &
          subset_dim(*), subset_bounds(2, *), pivots(*), pivot_order(*), pivot_shell_pairs(*), &
          pivot_shell_pair_loc(*), pivot_shell_pair_dim(*), pivot_orb_pairs(*)
    end function
    integer(c_int) function gi_chol_rkpq(ctx, dst) bind(C, name="gi_chol_rkpq")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dst(*)
    end function
    integer(c_int) function gi_chol_inv_l(ctx, dst) bind(C, name="gi_chol_inv_l")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dst(*)
    end function
    integer(c_int) function gi_chol_vdiag(ctx, dst) bind(C, name="gi_chol_vdiag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dst(*)
    end function
    integer(c_int) function gi_chol_timings(ctx, t) bind(C, name="gi_chol_timings")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: t(18)
    end function
    integer(c_int) function gi_chol_rkpq_device(ctx, ptr) bind(C, name="gi_chol_rkpq_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: ptr
    end function
    integer(c_int) function gi_eri_class_stats(ctx, cap, rows, n) bind(C, name="gi_eri_class_stats")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: cap
      real(c_double), intent(out) :: rows(5, *)
      integer(c_int64_t), intent(out) :: n
    end function
    integer(c_int) function gi_mo_transform_device(ctx, CA, lda, a0, a1, CB, ldb, b0, b1, max_buffer_mb, flags, &
        Rkab) bind(C, name="gi_mo_transform_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, CA, CB, Rkab
      integer(c_int64_t), value :: lda, a0, a1, ldb, b0, b1, max_buffer_mb, flags
    end function
    integer(c_int) function gi_ao_map(ctx, ordering, map, phase) bind(C, name="gi_ao_map")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ordering
      integer(c_int64_t), intent(out) :: map(*)
      real(c_double), intent(out) :: phase(*)
    end function
    integer(c_int) function gi_ao_rows_to_auto2e_device(ctx, ordering, X_ext, ld_ext, ncols, X_ao, ld_ao) &
        bind(C, name="gi_ao_rows_to_auto2e_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, X_ext, X_ao
      integer(c_int64_t), value :: ordering, ld_ext, ncols, ld_ao
    end function
    ! allocation-free variant (map and signs in the library's static device array; NAO <= 65536)
    integer(c_int) function gi_ao_rows_to_auto2e_device_noalloc(ctx, ordering, X_ext, ld_ext, ncols, X_ao, ld_ao) &
        bind(C, name="gi_ao_rows_to_auto2e_device_noalloc")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, X_ext, X_ao
      integer(c_int64_t), value :: ordering, ld_ext, ncols, ld_ao
    end function
    integer(c_int) function gi_mo_transform(ctx, CA, lda, a0, a1, CB, ldb, b0, b1, max_buffer_mb, flags, Rkab) &
        bind(C, name="gi_mo_transform")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: lda, a0, a1, ldb, b0, b1, max_buffer_mb, flags
      real(c_double), intent(in) :: CA(lda, *), CB(ldb, *)
      real(c_double), intent(out) :: Rkab(*)
    end function
    ! tensor hypercontraction: orbitals on caller-given grid points, the
    ! least-squares Z of the context's Cholesky vectors, MO-basis vectors
    integer(c_int) function gi_thc_grid_pivots(ctx, npts, xyz, qr_thresh, qr_thresh_reduced, block_dim, &
        npivots, npivots_reduced, pivots) bind(C, name="gi_thc_grid_pivots")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npts, block_dim
      real(c_double), intent(in) :: xyz(3, *)
      real(c_double), value :: qr_thresh, qr_thresh_reduced
      integer(c_int64_t), intent(out) :: npivots, npivots_reduced, pivots(*)
    end function
    integer(c_int) function gi_thc_orbitals(ctx, npts, xyz, normalize, Xgp, ldx) bind(C, name="gi_thc_orbitals")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npts, normalize, ldx
      real(c_double), intent(in) :: xyz(3, *)
      real(c_double), intent(out) :: Xgp(ldx, *)
    end function
    integer(c_int) function gi_thc_orbitals_device(ctx, npts, xyz, normalize, Xgp, ldx) &
        bind(C, name="gi_thc_orbitals_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, xyz, Xgp
      integer(c_int64_t), value :: npts, normalize, ldx
    end function
    integer(c_int) function gi_thc_z(ctx, npts, Xgp, ldx, npts_reduced, Zgk, ldz, Zgk_reduced, ldzr) &
        bind(C, name="gi_thc_z")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, Zgk_reduced
      integer(c_int64_t), value :: npts, ldx, npts_reduced, ldz, ldzr
      real(c_double), intent(in) :: Xgp(ldx, *)
      real(c_double), intent(out) :: Zgk(ldz, *)
    end function
    integer(c_int) function gi_thc_z_device(ctx, npts, Xgp, ldx, npts_reduced, Zgk, ldz, Zgk_reduced, ldzr) &
        bind(C, name="gi_thc_z_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, Xgp, Zgk, Zgk_reduced
      integer(c_int64_t), value :: npts, ldx, npts_reduced, ldz, ldzr
    end function
    integer(c_int) function gi_thc_rkab(ctx, npts, Xgp, ldx, Zgk, ldz, nvec, CA, lda, a0, a1, CB, ldb, b0, b1, &
        ordering, Rkab) bind(C, name="gi_thc_rkab")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npts, ldx, ldz, nvec, lda, a0, a1, ldb, b0, b1, ordering
      real(c_double), intent(in) :: Xgp(ldx, *), Zgk(ldz, *), CA(lda, *), CB(ldb, *)
      real(c_double), intent(out) :: Rkab(*)
    end function
    integer(c_int) function gi_thc_rkab_device(ctx, npts, Xgp, ldx, Zgk, ldz, nvec, CA, lda, a0, a1, CB, ldb, &
        b0, b1, ordering, Rkab) bind(C, name="gi_thc_rkab_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, Xgp, Zgk, CA, CB, Rkab
      integer(c_int64_t), value :: npts, ldx, ldz, nvec, lda, a0, a1, ldb, b0, b1, ordering
    end function
  end interface

contains

  ! message of the last failure as a Fortran string
  function gigpu_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
      end function
    end interface
    p = gi_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  subroutine gigpu_check(status)
    integer(c_int), intent(in) :: status
    if (status /= 0) then
      write(*, '(A)') "ERROR: " // gigpu_error_message()
      error stop
    end if
  end subroutine
end module gigpu_c


module gigpu_cholesky
  use iso_c_binding
  use gigpu_c
  implicit none

  ! TChol2Vecs (TwoStepCholesky_definitions.f90:56-74: dimensions, index
  ! arrays and Inv_L) and Rkpq (TCholeskyVecsOTF, Cholesky_Gammcor.f90:19-22)
  ! of one decomposition
  type gigpu_chol_vecs
    type(c_ptr) :: ctx = c_null_ptr
    integer(c_int64_t) :: NVecs = 0, NOrbPairs = 0, NShellPairs = 0, NSubsets = 0, MaxSubsetDim = 0
    integer(c_int64_t) :: NPivotShellPairs = 0
    integer(c_int64_t), allocatable :: ShellPairs(:, :), ShellPairLoc(:, :), ShellPairDim(:)
    integer(c_int64_t), allocatable :: SubsetDim(:), SubsetBounds(:, :), Pivots(:)
    integer(c_int64_t), allocatable :: PivotShellPairs(:), PivotShellPairLoc(:), PivotShellPairDim(:)
    integer(c_int64_t), allocatable :: PivotOrbPairs(:)
    real(c_double), allocatable :: Inv_L(:, :)
    real(c_double), allocatable :: Rkpq(:, :, :)
  end type

contains

  ! chol_gammcor_Rkpq over the TAOBasis components: Accuracy is one of
  ! GI_CHOL_ACCURACY_*; omega > 0 selects the range-separated operator
  ! (Kappa = 1/Omega^2, Cholesky_Gammcor.f90:411-415), omega <= 0 the full one.
  subroutine gigpu_chol_Rkpq(vecs, device, natoms, coords, nshells, shell_centers, shell_params, nparams, &
      momentum, nprim, max_nprim, cntr, exps, max_nang, norm, spher, accuracy, omega)
    type(gigpu_chol_vecs), intent(inout) :: vecs
    integer(c_int64_t), intent(in) :: device, natoms, nshells, nparams, max_nprim, max_nang, spher, accuracy
    real(c_double), intent(in) :: coords(3, natoms), cntr(max_nprim, nparams), exps(max_nprim, nparams)
    real(c_double), intent(in) :: norm(max_nang, nparams), omega
    integer(c_int64_t), intent(in) :: shell_centers(nshells), shell_params(nshells), momentum(nparams)
    integer(c_int64_t), intent(in) :: nprim(nparams)
    real(c_double) :: tau, kappa
    integer(c_int64_t) :: dims(8)
    integer(c_int64_t), allocatable :: pivot_order(:)

    if (.not. c_associated(vecs%ctx)) call gigpu_check(gi_create(vecs%ctx, device))
    call gigpu_check(gi_chol_accuracy_tau(accuracy, tau))
    kappa = 0.0_c_double
    if (omega > 0.0_c_double) kappa = 1.0_c_double / omega**2
    call gigpu_check(gi_set_basis(vecs%ctx, natoms, coords, nshells, shell_centers, shell_params, nparams, &
        momentum, nprim, max_nprim, cntr, exps, max_nang, norm, spher))
    call gigpu_check(gi_chol_build(vecs%ctx, tau, kappa, 4000_c_int64_t, 500_c_int64_t, 0_c_int64_t))
    call gigpu_check(gi_chol_dims(vecs%ctx, dims))
    vecs%NVecs = dims(1)
    vecs%NOrbPairs = dims(2)
    vecs%NShellPairs = dims(3)
    vecs%NSubsets = dims(4)
    vecs%MaxSubsetDim = dims(5)
    vecs%NPivotShellPairs = dims(6)
    if (allocated(vecs%ShellPairs)) deallocate(vecs%ShellPairs, vecs%ShellPairLoc, vecs%ShellPairDim, &
        vecs%SubsetDim, vecs%SubsetBounds, vecs%Pivots, vecs%PivotShellPairs, vecs%PivotShellPairLoc, &
        vecs%PivotShellPairDim, vecs%PivotOrbPairs, vecs%Inv_L, vecs%Rkpq)
    allocate(vecs%ShellPairs(2, vecs%NShellPairs), vecs%ShellPairLoc(3, vecs%NShellPairs), &
        vecs%ShellPairDim(vecs%NShellPairs), vecs%SubsetDim(vecs%NSubsets), vecs%SubsetBounds(2, vecs%NSubsets), &
        vecs%Pivots(vecs%NOrbPairs), vecs%PivotShellPairs(vecs%NPivotShellPairs), &
        vecs%PivotShellPairLoc(vecs%NPivotShellPairs), vecs%PivotShellPairDim(vecs%NPivotShellPairs), &
        vecs%PivotOrbPairs(vecs%NVecs), pivot_order(max(vecs%NVecs, 1_c_int64_t)))
    call gigpu_check(gi_chol_meta(vecs%ctx, vecs%ShellPairs, vecs%ShellPairLoc, vecs%ShellPairDim, &
        vecs%SubsetDim, vecs%SubsetBounds, vecs%Pivots, pivot_order, vecs%PivotShellPairs, &
        vecs%PivotShellPairLoc, vecs%PivotShellPairDim, vecs%PivotOrbPairs))
    allocate(vecs%Inv_L(vecs%NVecs, vecs%NVecs))
    call gigpu_check(gi_chol_inv_l(vecs%ctx, vecs%Inv_L))
    allocate(vecs%Rkpq(vecs%NVecs, vecs%MaxSubsetDim, vecs%NSubsets))
    call gigpu_check(gi_chol_rkpq(vecs%ctx, vecs%Rkpq))
  end subroutine

  ! chol_gammcor_Rkab: Rkab(NVecs, (a1-a0+1)*(b1-b0+1)), CA/CB rows already in
  ! the Auto2e AO order (Cholesky_Gammcor.f90:74-83 permutes them first)
  subroutine gigpu_chol_Rkab(Rkab, CA, a0, a1, CB, b0, b1, max_buffer_dim_mb, vecs)
    real(c_double), intent(out) :: Rkab(:, :)
    real(c_double), intent(in) :: CA(:, :), CB(:, :)
    integer(c_int64_t), intent(in) :: a0, a1, b0, b1, max_buffer_dim_mb
    type(gigpu_chol_vecs), intent(in) :: vecs
    if (size(Rkab, 1) /= vecs%NVecs .or. size(Rkab, 2) /= (a1 - a0 + 1) * (b1 - b0 + 1)) then
      write(*, '(A)') "ERROR: Invalid dimensions of Rkab"
      error stop
    end if
    call gigpu_check(gi_mo_transform(vecs%ctx, CA, int(size(CA, 1), c_int64_t), a0, a1, CB, &
        int(size(CB, 1), c_int64_t), b0, b1, max_buffer_dim_mb, 0_c_int64_t, Rkab))
  end subroutine

  subroutine gigpu_chol_free(vecs)
    type(gigpu_chol_vecs), intent(inout) :: vecs
    if (c_associated(vecs%ctx)) call gigpu_check(gi_destroy(vecs%ctx))
    vecs%ctx = c_null_ptr
  end subroutine
end module gigpu_cholesky
