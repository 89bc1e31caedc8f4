! Drives the THC entry points from Fortran through ISO_C_BINDING (module
! gigpu_c): prunes a parent grid (gi_thc_grid_pivots, thc_Chol_Pivots,
! TensorHypercontraction.f90:772-968), decomposes the Coulomb matrix
! (gi_chol_build), evaluates the orbitals on the pruned points
! (gi_thc_orbitals, thc_Grid :385-387), fits Z (gi_thc_z, thc_Z :393-476) and
! builds Rkab (gi_thc_rkab, thc_gammcor_Rkab, THC_Gammcor.f90:25-80), then
! writes everything to a stream file for the Python test to compare with the
! ctypes path.
!
!   thc_check input.bin output.bin
!
! input.bin (int64 / float64, stream): natoms nshells nparams max_nprim
! max_nang spher nparent block_dim nmo a0 a1 b0 b1 ordering, then tau kappa
! qr_thresh qr_thresh_reduced, coords(3, natoms), shell_centers,
! shell_params, momentum, nprim, cntr(max_nprim, nparams), exps(max_nprim,
! nparams), norm(max_nang, nparams), parent(3, nparent), CA(nao, nmo),
! CB(nao, nmo) with nao from the basis.
!
! output.bin: npivots npivots_reduced nvecs, pivots(npivots), then
! Xgp(npivots, nao), Zgk(npivots, nvecs), ZgkReduced(npivots_reduced, nvecs),
! Rkab(nvecs, NA*NB).
program thc_check
  use iso_c_binding
  use gigpu_c
  implicit none
  integer(c_int64_t) :: natoms, nshells, nparams, max_nprim, max_nang, spher, nparent, block_dim
  integer(c_int64_t) :: nmo, a0, a1, b0, b1, ordering, nao, npiv, nred, nvec, g
  integer(c_int64_t) :: dims(8)
  real(c_double) :: tau, kappa, qr, qr_red
  real(c_double), allocatable :: coords(:, :), cntr(:, :), exps(:, :), norm(:, :), parent(:, :)
  real(c_double), allocatable :: CA(:, :), CB(:, :), pts(:, :), Xgp(:, :), Zgk(:, :), Rkab(:, :)
  real(c_double), allocatable, target :: Zred(:, :)
  integer(c_int64_t), allocatable :: shell_centers(:), shell_params(:), momentum(:), nprim(:), piv(:)
  type(c_ptr) :: ctx
  character(len=4096) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old", action="read")
  read(u) natoms, nshells, nparams, max_nprim, max_nang, spher, nparent, block_dim, nmo, a0, a1, b0, b1, ordering
  read(u) tau, kappa, qr, qr_red
  allocate(coords(3, natoms), shell_centers(nshells), shell_params(nshells), momentum(nparams), nprim(nparams))
  allocate(cntr(max_nprim, nparams), exps(max_nprim, nparams), norm(max_nang, nparams), parent(3, nparent))
  read(u) coords, shell_centers, shell_params, momentum, nprim, cntr, exps, norm, parent

  call gigpu_check(gi_create(ctx, 0_c_int64_t))
  call gigpu_check(gi_set_basis(ctx, natoms, coords, nshells, shell_centers, shell_params, nparams, momentum, &
      nprim, max_nprim, cntr, exps, max_nang, norm, spher))
  nao = 0
  do g = 1, nshells
    if (spher /= 0) then
      nao = nao + 2 * momentum(shell_params(g)) + 1
    else
      nao = nao + ((momentum(shell_params(g)) + 1) * (momentum(shell_params(g)) + 2)) / 2
    end if
  end do
  allocate(CA(nao, nmo), CB(nao, nmo))
  read(u) CA, CB
  close(u)

  allocate(piv(nparent))
  call gigpu_check(gi_thc_grid_pivots(ctx, nparent, parent, qr, qr_red, block_dim, npiv, nred, piv))
  allocate(pts(3, npiv), Xgp(npiv, nao))
  do g = 1, npiv
    pts(:, g) = parent(:, piv(g) + 1)
  end do
  call gigpu_check(gi_thc_orbitals(ctx, npiv, pts, 1_c_int64_t, Xgp, npiv))
  call gigpu_check(gi_chol_build(ctx, tau, kappa, 4000_c_int64_t, 500_c_int64_t, 0_c_int64_t))
  call gigpu_check(gi_chol_dims(ctx, dims))
  nvec = dims(1)
  allocate(Zgk(npiv, nvec), Zred(max(nred, 1_c_int64_t), nvec))
  if (nred < npiv) then
    call gigpu_check(gi_thc_z(ctx, npiv, Xgp, npiv, nred, Zgk, npiv, c_loc(Zred), nred))
  else
    call gigpu_check(gi_thc_z(ctx, npiv, Xgp, npiv, 0_c_int64_t, Zgk, npiv, c_null_ptr, 1_c_int64_t))
  end if
  allocate(Rkab(nvec, (a1 - a0 + 1) * (b1 - b0 + 1)))
  call gigpu_check(gi_thc_rkab(ctx, npiv, Xgp, npiv, Zgk, npiv, nvec, CA, nao, a0, a1, CB, nao, b0, b1, ordering, &
      Rkab))
  call gigpu_check(gi_destroy(ctx))

  open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace", action="write")
  write(u) npiv, nred, nvec
  write(u) piv(1:npiv)
  write(u) Xgp, Zgk
  if (nred < npiv) write(u) Zred
  write(u) Rkab
  close(u)
  write(*, '(A,I0,A,I0)') "fortran thc ok: NGrid = ", npiv, " NVecs = ", nvec
end program
