! dsm_sgt_api.f90 -- iso_c_binding interface to include/dsm_sgt.h.
module dsm_sgt_api
  use iso_c_binding
  implicit none
  integer, parameter :: DSM_SGT_MAXZONE = 64
  integer(c_int32_t), parameter :: DSM_SGT_SH = 1, DSM_SGT_PSV = 2
  integer(c_int32_t), parameter :: DSM_SGT_FP64 = 0, DSM_SGT_FP32 = 1, DSM_SGT_BF16X3 = 2, DSM_SGT_BF16X6 = 3

  ! dsm_sgt_input_t
  type, bind(c) :: dsm_sgt_input
     integer(c_int32_t) :: nzone, maxlmax, imin, imax
     integer(c_int32_t) :: rsgtswitch, tsgtswitch, synnswitch, psgtswitch
     real(c_double) :: vrmin(DSM_SGT_MAXZONE), vrmax(DSM_SGT_MAXZONE)
     real(c_double) :: rho(4, DSM_SGT_MAXZONE), vpv(4, DSM_SGT_MAXZONE), vph(4, DSM_SGT_MAXZONE)
     real(c_double) :: vsv(4, DSM_SGT_MAXZONE), vsh(4, DSM_SGT_MAXZONE), eta(4, DSM_SGT_MAXZONE)
     real(c_double) :: qmu(DSM_SGT_MAXZONE), qkappa(DSM_SGT_MAXZONE)
     real(c_double) :: re, ratc, ratl, omegai_seed, tlen
     real(c_double) :: rmin_, rmax_, rdelta_, r0min
     real(c_double) :: thetamin, thetamax, thetadelta
     ! receiver lists in place of the grids (DSM3D_forward's files); 0: the grids
     integer(c_int32_t) :: r_list_n = 0, theta_list_n = 0
     type(c_ptr) :: r_list = c_null_ptr, theta_list = c_null_ptr
  end type dsm_sgt_input

  ! dsm_sgt_tuning_t (launch plan; 0 = default)
  type, bind(c) :: dsm_sgt_tuning
     integer(c_int32_t) :: nslot = 0, lchunk = 0, xcd_group = 0, graph = 0, catalog_window = 0
     integer(c_int32_t) :: overlap = 0
     integer(c_int32_t) :: reserved(2) = 0
  end type dsm_sgt_tuning

  ! dsm_km_params_t (include/dsm_km.h): KernelMaker's parameters of the catalog stage
  type, bind(c) :: dsm_km_params
     integer(c_int32_t) :: fmin, fmax, compo, reserved = 0
     real(c_double) :: c_red_reci, ipdistance
     real(c_double) :: mt(6)
  end type dsm_km_params

  interface
     integer(c_int32_t) function dsm_sgt_create(inp, mode, precision, device, handle) bind(c)
       import :: c_int32_t, c_ptr, dsm_sgt_input
       type(dsm_sgt_input), intent(in) :: inp
       integer(c_int32_t), value :: mode, precision, device
       type(c_ptr), intent(out) :: handle
     end function dsm_sgt_create
     integer(c_int32_t) function dsm_sgt_create_tuned(inp, mode, precision, device, plan, handle) bind(c)
       import :: c_int32_t, c_ptr, dsm_sgt_input, dsm_sgt_tuning
       type(dsm_sgt_input), intent(in) :: inp
       integer(c_int32_t), value :: mode, precision, device
       type(dsm_sgt_tuning), intent(in) :: plan
       type(c_ptr), intent(out) :: handle
     end function dsm_sgt_create_tuned
     integer(c_int32_t) function dsm_sgt_autotune(inp, mode, precision, device, cache, plan, probe_ms) bind(c)
       import :: c_int32_t, c_char, c_double, dsm_sgt_input, dsm_sgt_tuning
       type(dsm_sgt_input), intent(in) :: inp
       integer(c_int32_t), value :: mode, precision, device
       character(kind=c_char), dimension(*), intent(in) :: cache
       type(dsm_sgt_tuning), intent(out) :: plan
       real(c_double), intent(out) :: probe_ms
     end function dsm_sgt_autotune
     integer(c_int32_t) function dsm_sgt_run(handle, i0, i1) bind(c)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: handle
       integer(c_int32_t), value :: i0, i1
     end function dsm_sgt_run
     integer(c_int32_t) function dsm_sgt_run_list(handle, idx, n) bind(c)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: handle
       integer(c_int32_t), intent(in) :: idx(*)
       integer(c_int32_t), value :: n
     end function dsm_sgt_run_list
     integer(c_int32_t) function dsm_sgt_write_catalogs_async(handle, i0, i1, outdir, model) bind(c)
       import :: c_int32_t, c_ptr, c_char
       type(c_ptr), value :: handle
       integer(c_int32_t), value :: i0, i1
       character(kind=c_char), dimension(*), intent(in) :: outdir, model
     end function dsm_sgt_write_catalogs_async
     integer(c_int32_t) function dsm_sgt_wait(handle) bind(c)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: handle
     end function dsm_sgt_wait
     ! KernelMaker's catalog stage for npoints scattering points at receiver
     ! radius index ir (0-based): h3, h4 complex(8) (6, fmin:fmax, npoints)
     integer(c_int32_t) function dsm_km_points(psv, sh, ir, params, npoints, deltar, deltas, trig, h3, h4) bind(c)
       import :: c_int32_t, c_ptr, c_double, c_double_complex, dsm_km_params
       type(c_ptr), value :: psv, sh
       integer(c_int32_t), value :: ir, npoints
       type(dsm_km_params), intent(in) :: params
       real(c_double), intent(in) :: deltar(*), deltas(*), trig(12, *)
       complex(c_double_complex), intent(out) :: h3(6, *), h4(6, *)
     end function dsm_km_points
     ! vectorFFT_double for nvec spectra ccvec(imin:imax, nvec) -> rvec(iws:iwe, nvec)
     integer(c_int32_t) function dsm_km_vectorfft(h, nvec, imin, imax, np1, ccvec, omegai, tlen, iws, iwe, rvec) &
          bind(c)
       import :: c_int32_t, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nvec, imin, imax, np1, iws, iwe
       complex(c_double_complex), intent(in) :: ccvec(*)
       real(c_double), value :: omegai, tlen
       real(c_double), intent(out) :: rvec(*)
     end function dsm_km_vectorfft
     integer(c_int32_t) function dsm_sgt_lcut(handle, i, llog) bind(c)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: handle
       integer(c_int32_t), value :: i
       integer(c_int32_t), intent(out) :: llog
     end function dsm_sgt_lcut
     integer(c_int32_t) function dsm_sgt_write_catalog(handle, i, outdir, model) bind(c)
       import :: c_int32_t, c_ptr, c_char
       type(c_ptr), value :: handle
       integer(c_int32_t), value :: i
       character(kind=c_char), dimension(*), intent(in) :: outdir, model
     end function dsm_sgt_write_catalog
     type(c_ptr) function dsm_sgt_last_error(handle) bind(c)
       import :: c_ptr
       type(c_ptr), value :: handle
     end function dsm_sgt_last_error
     subroutine dsm_sgt_destroy(handle) bind(c)
       import :: c_ptr
       type(c_ptr), value :: handle
     end subroutine dsm_sgt_destroy
     integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! last error message of a handle as a Fortran string
  function dsm_sgt_message(handle) result(msg)
    type(c_ptr), intent(in) :: handle
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: s
    integer :: n, k
    s = dsm_sgt_last_error(handle)
    n = int(c_strlen(s))
    call c_f_pointer(s, p, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
       msg(k:k) = p(k)
    enddo
  end function dsm_sgt_message

end module dsm_sgt_api
