! sgtpsv_main.f90 -- `sgtpsv_gpu < database.inf`: drop-in for the SGTpsv
! driver (src/SGTpsv/SGTpsv-MPI.f90) on one GPU.
program sgtpsv_gpu
  use sgt_host_io
  implicit none
  call run_host(DSM_SGT_PSV)
end program sgtpsv_gpu
