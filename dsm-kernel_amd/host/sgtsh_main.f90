! sgtsh_main.f90 -- `sgtsh_gpu < database.inf`: drop-in for the SGTsh
! driver (src/SGTsh/SGTsh-MPI.f90) on one GPU.
program sgtsh_gpu
  use sgt_host_io
  implicit none
  call run_host(DSM_SGT_SH)
end program sgtsh_gpu
