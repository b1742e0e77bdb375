! km_harness.f90 -- C entry points around the reference's KernelMaker catalog
! routines (src/KernelMaker/fourier.f90: rdsgtomega, clsgt, rsgt2h3freq,
! tsgt2h4freq, vectorFFT_double), compiled with the reference's own module.f90
! and fourier.f90 into oracle/_ref/libref_km.so (oracle/Makefile).
!
! TEST INFRASTRUCTURE ONLY: it sets the module variables those routines read
! (as KernelMaker-1.0.0.f90 and its input reader do), calls them unchanged,
! and copies their results out, so that tests can check the C oracle
! (km_oracle.c) against the reference's arithmetic.
module km_harness
  use iso_c_binding
  use parameters
  use angles
  use tmpSGTs
  use kernels
  implicit none
contains

  subroutine c2f(cs, fs)
    character(kind=c_char), intent(in) :: cs(*)
    character(len=*), intent(out) :: fs
    integer :: k
    fs = ' '
    do k = 1, len(fs)
       if (cs(k) == c_null_char) exit
       fs(k:k) = cs(k)
    enddo
  end subroutine c2f

  ! catalog grid, frequencies, reduction slowness, interpolation window,
  ! moment tensor, component (0 Z, 1 R, 2 T), catalog directory and model
  subroutine km_ref_init(ntheta_, thmin, thdelta, fmin_, fmax_, tlen_, cred, ipd, mt_, compo_, outdir, model) &
       bind(c, name='km_ref_init')
    integer(c_int), value :: ntheta_, fmin_, fmax_, compo_
    real(c_double), value :: thmin, thdelta, tlen_, cred, ipd
    real(c_double), intent(in) :: mt_(6)
    character(kind=c_char), intent(in) :: outdir(*), model(*)
    integer :: i
    theta_n = ntheta_
    thetamin = thmin
    thetadelta = thdelta
    fmin = fmin_
    fmax = fmax_
    tlen = tlen_
    c_red_reci = cred
    ipdistance = ipd
    mt(1:6) = mt_(1:6)
    compo = 'Z'
    if (compo_ == 1) compo = 'R'
    if (compo_ == 2) compo = 'T'
    call c2f(outdir, PoutputDir)
    call c2f(model, modelname)
    if (allocated(thetaD)) deallocate(thetaD)
    allocate(thetaD(1:theta_n))
    do i = 1, theta_n
       thetaD(i) = (thetamin + dble(i-1)*thetadelta)
    enddo
    if (allocated(tsgtomega)) deallocate(tsgtomega, rsgtomega, synnomega, tsgtF, rsgtF, synnF, h3, h4)
    allocate(tsgtomega(1:20, fmin:fmax, 1:theta_n), rsgtomega(1:10, fmin:fmax, 1:theta_n))
    allocate(synnomega(1:10, fmin:fmax, 1:theta_n))
    allocate(tsgtF(1:20, fmin:fmax), rsgtF(1:10, fmin:fmax), synnF(1:10, fmin:fmax))
    allocate(h3(1:6, fmin:fmax), h4(1:6, fmin:fmax))
    tsgtomega = cmplx(0.d0); rsgtomega = cmplx(0.d0); synnomega = cmplx(0.d0)
    h3 = cmplx(0.d0); h4 = cmplx(0.d0)
    if (allocated(tmph01)) deallocate(tmph01, tmph02, tmph03, tmph04, tmph05, tmph06, tmph07, tmph08, tmph09, &
         tmph10, tmph11, tmph12, tmph13, tmph14, tmph15, tmph16, tmph17, tmph18, tmph19, tmph20, tmph21, tmph22, &
         tmph23, tmph24, tmph25, tmph26)
    allocate(tmph01(fmin:fmax), tmph02(fmin:fmax), tmph03(fmin:fmax), tmph04(fmin:fmax), tmph05(fmin:fmax), &
         tmph06(fmin:fmax), tmph07(fmin:fmax), tmph08(fmin:fmax), tmph09(fmin:fmax), tmph10(fmin:fmax), &
         tmph11(fmin:fmax), tmph12(fmin:fmax), tmph13(fmin:fmax), tmph14(fmin:fmax), tmph15(fmin:fmax), &
         tmph16(fmin:fmax), tmph17(fmin:fmax), tmph18(fmin:fmax), tmph19(fmin:fmax), tmph20(fmin:fmax), &
         tmph21(fmin:fmax), tmph22(fmin:fmax), tmph23(fmin:fmax), tmph24(fmin:fmax), tmph25(fmin:fmax), &
         tmph26(fmin:fmax))
    if (allocated(crq)) deallocate(crq, srq, crq2, srq2, csq, ssq, csq2, ssq2, cqs, sqs, cqs2, sqs2)
    allocate(crq(0:1,0:1), srq(0:1,0:1), crq2(0:1,0:1), srq2(0:1,0:1), csq(0:1,0:1), ssq(0:1,0:1))
    allocate(csq2(0:1,0:1), ssq2(0:1,0:1), cqs(0:1,0:1), sqs(0:1,0:1), cqs2(0:1,0:1), sqs2(0:1,0:1))
  end subroutine km_ref_init

  ! rdsgtomega for one file kind (ipsvorsh 2/1 SYNN, 20/10 RSGT, 200/100 TSGT)
  subroutine km_ref_read(rx_, ry_, num_sgt, num_psv, ipsvorsh) bind(c, name='km_ref_read')
    real(c_double), value :: rx_, ry_
    integer(c_int), value :: num_sgt, num_psv, ipsvorsh
    call rdsgtomega(rx_, ry_, num_sgt, num_psv, ipsvorsh)
  end subroutine km_ref_read

  ! the catalog rdsgtomega built: kind 0 TSGT (20), 1 RSGT (10), 2 SYNN (10),
  ! out(num, fmin:fmax, theta_n)
  subroutine km_ref_omega(kind, out) bind(c, name='km_ref_omega')
    integer(c_int), value :: kind
    complex(c_double_complex), intent(out) :: out(*)
    integer :: n
    if (kind == 0) then
       n = size(tsgtomega)
       out(1:n) = reshape(tsgtomega, [n])
    elseif (kind == 1) then
       n = size(rsgtomega)
       out(1:n) = reshape(rsgtomega, [n])
    else
       n = size(synnomega)
       out(1:n) = reshape(synnomega, [n])
    endif
  end subroutine km_ref_omega

  ! one scattering point as calculateKernel (calculateK.f90:131-162) takes
  ! it: clsgt of RSGT at deltar and of TSGT at deltas, rsgt2h3freq,
  ! tsgt2h4freq.  trig: crq srq crq2 srq2 csq ssq csq2 ssq2 cqs sqs cqs2 sqs2.
  ! h3out, h4out: (6, fmin:fmax)
  subroutine km_ref_point(dr, ds, trig, h3out, h4out) bind(c, name='km_ref_point')
    real(c_double), value :: dr, ds
    real(c_double), intent(in) :: trig(12)
    complex(c_double_complex), intent(out) :: h3out(*), h4out(*)
    integer :: n
    ip = 1
    ith = 1
    crq(1,1) = trig(1); srq(1,1) = trig(2); crq2(1,1) = trig(3); srq2(1,1) = trig(4)
    csq(1,1) = trig(5); ssq(1,1) = trig(6); csq2(1,1) = trig(7); ssq2(1,1) = trig(8)
    cqs(1,1) = trig(9); sqs(1,1) = trig(10); cqs2(1,1) = trig(11); sqs2(1,1) = trig(12)
    rsgtF = dcmplx(0.d0)
    tsgtF = dcmplx(0.d0)
    call clsgt(dr, 10, rsgtF(1:10, fmin:fmax), rsgtomega(1:10, fmin:fmax, 1:theta_n))
    call clsgt(ds, 20, tsgtF(1:20, fmin:fmax), tsgtomega(1:20, fmin:fmax, 1:theta_n))
    call rsgt2h3freq
    call tsgt2h4freq
    n = size(h3)
    h3out(1:n) = reshape(h3, [n])
    h4out(1:n) = reshape(h4, [n])
  end subroutine km_ref_point

  ! vectorFFT_double as the reference calls it
  subroutine km_ref_vectorfft(imin_, imax_, np1_, ccvec, rvec, omegai_, tlen_, iws, iwe) &
       bind(c, name='km_ref_vectorfft')
    integer(c_int), value :: imin_, imax_, np1_, iws, iwe
    complex(c_double_complex), intent(in) :: ccvec(imin_:imax_)
    real(c_double), intent(out) :: rvec(iws:iwe)
    real(c_double), value :: omegai_, tlen_
    complex(kind(0d0)) :: cc(imin_:imax_)
    cc = ccvec
    call vectorFFT_double(imin_, imax_, np1_, cc, rvec, omegai_, tlen_, iws, iwe)
  end subroutine km_ref_vectorfft

end module km_harness
