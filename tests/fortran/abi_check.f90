! Host-only check of the ISO_C_BINDING interface: accuracy levels map to the
! reference thresholds (Cholesky_Gammcor.f90:398-410) and an invalid level
! comes back as a status with the reference's message.  No GPU call.
program abi_check
  use iso_c_binding
  use gigpu_c
  implicit none
  real(c_double) :: tau
  integer(c_int) :: st
  call gigpu_check(gi_chol_accuracy_tau(GI_CHOL_ACCURACY_DEFAULT, tau))
  if (tau /= 1.0e-5_c_double) error stop 1
  call gigpu_check(gi_chol_accuracy_tau(GI_CHOL_ACCURACY_TIGHT, tau))
  if (tau /= 1.0e-6_c_double) error stop 2
  call gigpu_check(gi_chol_accuracy_tau(GI_CHOL_ACCURACY_LUDICROUS, tau))
  if (tau /= 1.0e-7_c_double) error stop 3
  call gigpu_check(gi_chol_accuracy_tau(GI_CHOL_ACCURACY_DEBUG, tau))
  if (tau /= 1.0e-10_c_double) error stop 4
  st = gi_chol_accuracy_tau(9_c_int64_t, tau)
  if (st == 0) error stop 5
  if (gigpu_error_message() /= "Invalid accuracy setting") error stop 6
  write(*, '(A)') "fortran abi ok"
end program
