! The per-event drop-in: a program written like simc.f's montecarlo() that
! calls mc_hms / mc_sos / mc_hrsr / mc_hrsl / mc_shms once per event, with the common blocks the reference
! declares for the STOP counters (/spectrometer1/ to /spectrometer4/, /spectrometer_shms/),
! the track (/track/) and ctau (/decd/), linked against libsimc_dropin.so.
! Reads n events (SoA float64 dpp, x, y, z, dxdz, dydz), writes the per-event
! outputs (SoA float64 dpp, y, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, pathlen,
! then 4-byte ok) and prints the counters and, for accepted events, the
! largest difference between mc_*_recon on /track/ and the recon of the arm.
!   usage: dropin_events ARM N P_SPEC M2 SEED IN_FILE OUT_FILE  (SIMC_MAP_DIR set)
program dropin_events
  implicit none
  integer*4 hstop(22), sstop(52), ostop(18), rstop(21), lstop(21)
  common /spectrometer1/ hstop
  common /spectrometer2/ ostop
  common /spectrometer3/ rstop
  common /spectrometer4/ lstop
  common /spectrometer_shms/ sstop
  real*8 trk(80)
  common /track/ trk
  real*8 ctau, decd_rest(40)
  common /decd/ ctau, decd_rest
  character(len=512) :: in_file, out_file, arg
  integer :: arm, u, i, k
  integer*8 :: n, seed
  real*8 :: p_spec, m2, m2e, th_spec, resmult, pathlen, fry
  real*8 :: dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp
  real*8 :: r_dpp, r_dth, r_dph, r_y, worst
  real*8, allocatable :: ev(:, :), o(:, :)
  integer*4, allocatable :: okv(:)
  logical :: ok_spec, ms_flag, wcs_flag, decay_flag, using_coll
  integer*4 :: spectr

  call get_command_argument(1, arg); read (arg, *) arm
  call get_command_argument(2, arg); read (arg, *) n
  call get_command_argument(3, arg); read (arg, *) p_spec
  call get_command_argument(4, arg); read (arg, *) m2
  call get_command_argument(5, arg); read (arg, *) seed
  call get_command_argument(6, in_file)
  call get_command_argument(7, out_file)
  allocate(ev(n, 6), o(n, 9), okv(n))
  open (newunit=u, file=trim(in_file), access='stream', form='unformatted', status='old')
  read (u) ev
  close (u)

  hstop = 0
  sstop = 0
  ostop = 0
  rstop = 0
  lstop = 0
  ctau = 0
  call simc_dropin_seed(seed)
  ms_flag = .true.
  wcs_flag = .true.
  decay_flag = .false.
  using_coll = .false.
  spectr = 5
  th_spec = 0
  worst = 0
  do i = 1, n
    dpp = ev(i, 1); x = ev(i, 2); y = ev(i, 3); z = ev(i, 4); dxdz = ev(i, 5); dydz = ev(i, 6)
    fry = x
    pathlen = 0
    x_fp = 0; dx_fp = 0; y_fp = 0; dy_fp = 0
    m2e = m2
    if (arm == 1) then
      call mc_hms(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2e, &
                  ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen, using_coll)
    else if (arm == 2) then
      call mc_sos(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2e, &
                  ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen)
    else if (arm == 3) then
      call mc_hrsr(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2e, &
                   ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen)
    else if (arm == 4) then
      call mc_hrsl(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2e, &
                   ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen)
    else
      call mc_shms(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, m2e, &
                   ms_flag, wcs_flag, decay_flag, resmult, fry, ok_spec, pathlen, spectr, using_coll)
    end if
    if (ok_spec) then
      if (arm == 1) then
        call mc_hms_recon(r_dpp, r_dth, r_dph, r_y, fry)
      else if (arm == 2) then
        call mc_sos_recon(r_dpp, r_dth, r_dph, r_y, fry)
      else if (arm == 3) then
        call mc_hrsr_recon(r_dpp, r_dth, r_dph, r_y, fry)
      else if (arm == 4) then
        call mc_hrsl_recon(r_dpp, r_dth, r_dph, r_y, fry)
      else
        call mc_shms_recon(r_dpp, r_dth, r_dph, r_y, fry, spectr)
      end if
      worst = max(worst, abs(r_dpp - dpp), abs(r_dph - dxdz), abs(r_dth - dydz), abs(r_y - y))
    end if
    o(i, :) = [dpp, y, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp, pathlen]
    okv(i) = 0
    if (ok_spec) okv(i) = 1
  end do
  call simc_dropin_close()

  write (*, '(a)', advance='no') 'counters'
  if (arm == 1) then
    do k = 1, 22
      write (*, '(1x,i0)', advance='no') hstop(k)
    end do
  else if (arm == 2) then
    do k = 1, 18
      write (*, '(1x,i0)', advance='no') ostop(k)
    end do
  else if (arm == 3) then
    do k = 1, 21
      write (*, '(1x,i0)', advance='no') rstop(k)
    end do
  else if (arm == 4) then
    do k = 1, 21
      write (*, '(1x,i0)', advance='no') lstop(k)
    end do
  else
    do k = 1, 52
      write (*, '(1x,i0)', advance='no') sstop(k)
    end do
  end if
  write (*, *)
  write (*, '(a,es12.4)') 'recon_worst ', worst
  open (newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
  write (u) o, okv
  close (u)
end program dropin_events
