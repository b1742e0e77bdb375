! sgt_host.f90 -- Fortran host of the MI355X SGT engine (module; the programs
! are sgtpsv_main.f90 and sgtsh_main.f90: `sgtpsv_gpu < database.inf`,
! `sgtsh_gpu < database.inf`).  Same stdin grammar, DSM.config and model files as the
! reference's SGTsh / SGTpsv drivers (src/SGTsh/others.f90:1-110), same
! catalog files and log files; the per-frequency work is done on one GPU
! through the C-ABI (include/dsm_sgt.h).  No MPI: frequencies are batched on
! the device instead of distributed over ranks.
module sgt_host_io
  use dsm_sgt_api
  implicit none
contains

  ! Copy a text file minus '#' lines up to a line starting 'end' into recs.
  subroutine read_records(unit, recs, n)
    integer, intent(in) :: unit
    character(len=120), intent(out) :: recs(:)
    integer, intent(out) :: n
    character(len=120) :: line
    integer :: ios
    n = 0
    do
       read(unit, '(a120)', iostat=ios) line
       if (ios /= 0) exit
       if (line(1:1) == '#') cycle
       if (line(1:3) == 'end') exit
       n = n + 1
       recs(n) = line
    enddo
  end subroutine read_records

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(0, '(a)') 'sgt_gpu: '//trim(msg)
    error stop 1
  end subroutine fail

  subroutine run_host(mode)
    integer(c_int32_t), intent(in) :: mode
    type(dsm_sgt_input) :: inp
    character(len=120) :: recs(200), crecs(50), mrecs(2000)
    character(len=120) :: dsmconf, outdir, modelf, modelname
    character(len=200) :: list, list1, fname
    character(len=23) :: tag
    character(len=8) :: datex
    character(len=10) :: timex
    character(len=10) :: kind
    integer :: n, nc, nm, ios, i, j, k, iz, sw(4), pos
    real(c_double) :: vals(28)
    type(c_ptr) :: h
    integer(c_int32_t) :: rc, llog, prec, i0, i1, win
    type(dsm_sgt_tuning) :: plan
    real(c_double) :: probe
    character(len=256) :: envv
    integer :: est

    call read_records(5, recs, n)
    if (n < 10) call fail('parameter file: expected 10 records')
    dsmconf = adjustl(recs(1)); outdir = adjustl(recs(2))
    modelf = adjustl(recs(3)); modelname = adjustl(recs(4))
    read(recs(5), *) inp%tlen
    read(recs(6), *) inp%rmin_, inp%rmax_, inp%rdelta_
    read(recs(7), *) inp%r0min
    read(recs(8), *) inp%thetamin, inp%thetamax, inp%thetadelta
    read(recs(9), *) inp%imin, inp%imax
    sw = 0
    read(recs(10), *, iostat=ios) sw(1:4)   ! the shipped examples give 3 switches
    if (ios /= 0) read(recs(10), *) sw(1:3)
    inp%rsgtswitch = sw(1); inp%tsgtswitch = sw(2); inp%synnswitch = sw(3); inp%psgtswitch = sw(4)
    call execute_command_line('mkdir -p '//trim(outdir)//'/RSGT '//trim(outdir)//'/TSGT '//trim(outdir)//'/log')

    open(11, file=dsmconf, status='old', action='read', iostat=ios)
    if (ios /= 0) call fail('cannot open '//trim(dsmconf))
    call read_records(11, crecs, nc)
    close(11)
    read(crecs(1), *) inp%re
    read(crecs(2), *) inp%ratc
    read(crecs(3), *) inp%ratl
    read(crecs(4), *) inp%omegai_seed
    read(crecs(5), *) inp%maxlmax

    open(12, file=modelf, status='old', action='read', iostat=ios)
    if (ios /= 0) call fail('cannot open '//trim(modelf))
    call read_records(12, mrecs, nm)
    close(12)
    read(mrecs(1), *) inp%nzone
    if (inp%nzone < 1 .or. inp%nzone > DSM_SGT_MAXZONE) call fail('nzone out of range')
    ! each zone: 28 list-directed values starting on a fresh record
    k = 2
    do iz = 1, inp%nzone
       pos = 0
       do while (pos < 28)
          if (k > nm) call fail('model file ends inside a zone')
          j = count_values(mrecs(k))
          read(mrecs(k), *) vals(pos+1:min(28, pos+j))
          pos = min(28, pos + j)
          k = k + 1
       enddo
       inp%vrmin(iz) = vals(1); inp%vrmax(iz) = vals(2)
       inp%rho(:, iz) = vals(3:6); inp%vpv(:, iz) = vals(7:10); inp%vph(:, iz) = vals(11:14)
       inp%vsv(:, iz) = vals(15:18); inp%vsh(:, iz) = vals(19:22); inp%eta(:, iz) = vals(23:26)
       inp%qmu(iz) = vals(27); inp%qkappa(iz) = vals(28)
    enddo

    kind = 'SH'
    if (mode == DSM_SGT_PSV) kind = ''
    write(tag, '(I7,".",I7)') inp%imin, inp%imax
    do j = 1, 15
       if (tag(j:j) == ' ') tag(j:j) = '0'
    enddo
    list = trim(outdir)//'/log/calLog'//trim(kind)//'.'//trim(modelname)//'.'//tag(1:15)
    open(21, file=list, status='unknown', form='formatted')
    call date_and_time(datex, timex)
    write(21, '(/a,a4,a,a2,a,a2,a,a2,a,a2,a,a4)') '    Starting date and time:                     ', &
         datex(1:4), '-', datex(5:6), '-', datex(7:8), '.  ', timex(1:2), ':', timex(3:4), ':', timex(5:8)
    close(21)
    if (mode == DSM_SGT_PSV) then   ! SGTpsv-MPI.f90:522-531
       open(21, file=list, status='old', access='append', form='formatted')
       call date_and_time(datex, timex)
       write(21, '(/a,a4,a,a2,a,a2,a,a2,a,a2,a,a4)') '    here we go!:                     ', &
            datex(1:4), '-', datex(5:6), '-', datex(7:8), '.  ', timex(1:2), ':', timex(3:4), ':', timex(5:8)
       close(21)
    endif
    write(tag, '(I7,".",I7,".",I7)') inp%imin, inp%imax, 0
    do j = 1, 22
       if (tag(j:j) == ' ') tag(j:j) = '0'
    enddo
    list1 = trim(outdir)//'/log/list'//trim(kind)//'.'//trim(modelname)//'.'//tag

    ! form: DSM_SGT_PRECISION = fp64 (default, the parity form) | fp32 |
    ! bf16x3 | bf16x6 (the batched mixed form)
    prec = DSM_SGT_FP64
    call get_environment_variable('DSM_SGT_PRECISION', envv, status=est)
    if (est == 0) then
       select case (trim(envv))
       case ('fp64'); prec = DSM_SGT_FP64
       case ('fp32'); prec = DSM_SGT_FP32
       case ('bf16x3'); prec = DSM_SGT_BF16X3
       case ('bf16x6'); prec = DSM_SGT_BF16X6
       case default; call fail('DSM_SGT_PRECISION: fp64, fp32, bf16x3 or bf16x6')
       end select
    endif
    ! launch plan: autotuned when DSM_SGT_TUNE_CACHE names a cache file;
    ! the device catalog holds DSM_SGT_WINDOW frequencies (default 128),
    ! written out window by window: each window's files are written by the
    ! library's writer thread while the device computes the next window
    call get_environment_variable('DSM_SGT_TUNE_CACHE', envv, status=est)
    if (est == 0 .and. len_trim(envv) > 0) then
       rc = dsm_sgt_autotune(inp, mode, prec, 0_c_int32_t, trim(envv)//c_null_char, plan, probe)
       if (rc /= 0) call fail('autotune failed')
    endif
    win = 128
    call get_environment_variable('DSM_SGT_WINDOW', envv, status=est)
    if (est == 0) read(envv, *) win
    if (win < 1) call fail('DSM_SGT_WINDOW must be positive')
    plan%catalog_window = win
    rc = dsm_sgt_create_tuned(inp, mode, prec, 0_c_int32_t, plan, h)
    if (rc /= 0) call fail(dsm_sgt_message(h))

    open(24, file=list1, status='unknown', form='formatted')
    write(24, *)
    i0 = inp%imin
    do while (i0 <= inp%imax)
       ! win computed frequencies (i = 0 is skipped)
       i1 = min(inp%imax, max(i0, 1_c_int32_t) + win - 1)
       rc = dsm_sgt_run(h, i0, i1)
       if (rc /= 0) call fail(dsm_sgt_message(h))
       do i = i0, i1
          if (i == 0) cycle
          rc = dsm_sgt_lcut(h, i, llog)
          if (rc /= 0) call fail(dsm_sgt_message(h))
          write(24, *) i, dble(i) / inp%tlen, llog
       enddo
       if (i1 >= 1) then
          rc = dsm_sgt_write_catalogs_async(h, i0, i1, trim(outdir)//c_null_char, trim(modelname)//c_null_char)
          if (rc /= 0) call fail(dsm_sgt_message(h))
       endif
       i0 = i1 + 1
    enddo
    rc = dsm_sgt_wait(h)
    if (rc /= 0) call fail(dsm_sgt_message(h))
    close(24)
    call dsm_sgt_destroy(h)

    open(21, file=list, status='old', access='append', form='formatted')
    call date_and_time(datex, timex)
    write(21, '(/a,a4,a,a2,a,a2,a,a2,a,a2,a,a4)') '    Finishing date and time:                     ', &
         datex(1:4), '-', datex(5:6), '-', datex(7:8), '.  ', timex(1:2), ':', timex(3:4), ':', timex(5:8)
    close(21)
    fname = ''
  end subroutine run_host

  ! number of blank/comma separated values on a record
  integer function count_values(rec)
    character(len=*), intent(in) :: rec
    integer :: k
    logical :: inval
    count_values = 0
    inval = .false.
    do k = 1, len_trim(rec)
       if (rec(k:k) == ' ' .or. rec(k:k) == ',' .or. rec(k:k) == char(9)) then
          inval = .false.
       else if (.not. inval) then
          inval = .true.
          count_values = count_values + 1
       endif
    enddo
  end function count_values

end module sgt_host_io
