! Batch driver through the Fortran module simc_gpu: reads n events (SoA
! float64: dpp, x, y, z, dxdz, dydz) from a raw file, transports them through
! one arm with simc_arm_run_host (Philox stream, given seed), writes the
! outputs (SoA float64: dpp, y, dxdz, dydz, x_fp, dx_fp, y_fp, dy_fp,
! pathlen, then uint8 ok) and prints the STOP counters.
!   usage: batch_host MAP_DIR ARM N P_SPEC M2 SEED IN_FILE OUT_FILE
program batch_host
  use iso_c_binding
  use simc_gpu
  implicit none
  character(len=512) :: map_dir, in_file, out_file, arg
  integer(c_int) :: arm, rc
  integer(c_int64_t) :: n, seed
  real(c_double) :: p_spec, m2
  real(c_double), allocatable, target :: ev(:, :), o(:, :)
  integer(c_int8_t), allocatable, target :: ok(:)
  integer(c_int64_t) :: cnt(SIMC_SHMS_NCOUNTERS)
  type(c_ptr) :: ctx
  type(simc_settings) :: s
  type(simc_events_in) :: evin
  type(simc_events_out) :: evout
  type(simc_rng) :: rng
  integer :: u, k, nc

  call get_command_argument(1, map_dir)
  call get_command_argument(2, arg); read (arg, *) arm
  call get_command_argument(3, arg); read (arg, *) n
  call get_command_argument(4, arg); read (arg, *) p_spec
  call get_command_argument(5, arg); read (arg, *) m2
  call get_command_argument(6, arg); read (arg, *) seed
  call get_command_argument(7, in_file)
  call get_command_argument(8, out_file)

  allocate(ev(n, 6), o(n, 9), ok(n))
  open (newunit=u, file=trim(in_file), access='stream', form='unformatted', status='old')
  read (u) ev
  close (u)

  rc = simc_ctx_create(trim(map_dir)//c_null_char, ctx)
  if (rc /= SIMC_OK) then
    print *, 'simc_ctx_create failed: ', simc_error_message()
    stop 1
  end if

  s%p_spec = p_spec
  s%m2 = m2
  s%fit_mode = SIMC_FIT_FP64
  o(:, 1) = ev(:, 1)   ! dpp   (INOUT)
  o(:, 2) = ev(:, 3)   ! y     (INOUT)
  o(:, 3) = ev(:, 5)   ! dxdz  (INOUT)
  o(:, 4) = ev(:, 6)   ! dydz  (INOUT)
  o(:, 5:9) = 0
  evin%dpp = c_loc(ev(1, 1)); evin%x = c_loc(ev(1, 2)); evin%y = c_loc(ev(1, 3))
  evin%z = c_loc(ev(1, 4)); evin%dxdz = c_loc(ev(1, 5)); evin%dydz = c_loc(ev(1, 6))
  evout%dpp = c_loc(o(1, 1)); evout%y = c_loc(o(1, 2)); evout%dxdz = c_loc(o(1, 3))
  evout%dydz = c_loc(o(1, 4)); evout%x_fp = c_loc(o(1, 5)); evout%dx_fp = c_loc(o(1, 6))
  evout%y_fp = c_loc(o(1, 7)); evout%dy_fp = c_loc(o(1, 8)); evout%pathlen = c_loc(o(1, 9))
  evout%ok = c_loc(ok(1))
  rng%kind = SIMC_RNG_PHILOX
  rng%seed = seed
  rng%event0 = 0

  rc = simc_arm_run_host(ctx, arm, s, evin, evout, n, rng)
  if (rc /= SIMC_OK) then
    print *, 'simc_arm_run_host failed: ', simc_error_message()
    stop 1
  end if

  nc = SIMC_HMS_NCOUNTERS
  if (arm == SIMC_ARM_SHMS) nc = SIMC_SHMS_NCOUNTERS
  if (arm == SIMC_ARM_SOS) nc = SIMC_SOS_NCOUNTERS
  if (arm == SIMC_ARM_HRSR .or. arm == SIMC_ARM_HRSL) nc = SIMC_HRS_NCOUNTERS
  rc = simc_counters_get(ctx, arm, cnt, nc)
  write (*, '(a)', advance='no') 'counters'
  do k = 1, nc
    write (*, '(1x,i0)', advance='no') cnt(k)
  end do
  write (*, *)

  open (newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
  write (u) o, ok
  close (u)
  call simc_ctx_destroy(ctx)
end program batch_host
