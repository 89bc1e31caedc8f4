! Drives the Fortran side of the boundary on the GPU: gigpu_chol_Rkpq and
! gigpu_chol_Rkab (the module gigpu_cholesky wrappers of chol_gammcor_Rkpq
! and chol_gammcor_Rkab, Cholesky_Gammcor.f90:26-92,335-377) on a basis read
! from a stream file, then writes the TChol2Vecs dimensions and index arrays,
! Inv_L, Rkpq and Rkab to a second stream file for the Python test to compare
! with the ctypes path.
!
!   chol_check input.bin output.bin
!
! input.bin (int64 / float64, stream): natoms nshells nparams max_nprim
! max_nang spher accuracy nao nmo a0 a1 b0 b1 max_buffer_mb, then coords(3,
! natoms), shell_centers, shell_params, momentum, nprim, cntr(max_nprim,
! nparams), exps(max_nprim, nparams), norm(max_nang, nparams), CA(nao, nmo),
! CB(nao, nmo).
program chol_check
  use iso_c_binding
  use gigpu_cholesky
  implicit none
  integer(c_int64_t) :: natoms, nshells, nparams, max_nprim, max_nang, spher, accuracy
  integer(c_int64_t) :: nao, nmo, a0, a1, b0, b1, max_buffer_mb
  real(c_double), allocatable :: coords(:, :), cntr(:, :), exps(:, :), norm(:, :), CA(:, :), CB(:, :)
  real(c_double), allocatable :: Rkab(:, :)
  integer(c_int64_t), allocatable :: shell_centers(:), shell_params(:), momentum(:), nprim(:)
  type(gigpu_chol_vecs) :: vecs
  character(len=4096) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old", action="read")
  read(u) natoms, nshells, nparams, max_nprim, max_nang, spher, accuracy, nao, nmo, a0, a1, b0, b1, max_buffer_mb
  allocate(coords(3, natoms), shell_centers(nshells), shell_params(nshells), momentum(nparams), nprim(nparams))
  allocate(cntr(max_nprim, nparams), exps(max_nprim, nparams), norm(max_nang, nparams), CA(nao, nmo), CB(nao, nmo))
  read(u) coords, shell_centers, shell_params, momentum, nprim, cntr, exps, norm, CA, CB
  close(u)

  call gigpu_chol_Rkpq(vecs, 0_c_int64_t, natoms, coords, nshells, shell_centers, shell_params, nparams, &
      momentum, nprim, max_nprim, cntr, exps, max_nang, norm, spher, accuracy, 0.0_c_double)
  allocate(Rkab(vecs%NVecs, (a1 - a0 + 1) * (b1 - b0 + 1)))
  call gigpu_chol_Rkab(Rkab, CA, a0, a1, CB, b0, b1, max_buffer_mb, vecs)

  open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace", action="write")
  write(u) vecs%NVecs, vecs%NOrbPairs, vecs%NShellPairs, vecs%NSubsets, vecs%MaxSubsetDim, vecs%NPivotShellPairs
  write(u) vecs%ShellPairs, vecs%ShellPairLoc, vecs%Pivots, vecs%PivotOrbPairs, vecs%Inv_L, vecs%Rkpq, Rkab
  close(u)
  call gigpu_chol_free(vecs)
  write(*, '(A,I0)') "fortran chol ok: NVecs = ", vecs%NVecs
end program
