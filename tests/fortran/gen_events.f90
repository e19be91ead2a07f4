! Generator driver through the Fortran module simc_gpu: H(e,e'p) elastic
! events with simc_gen_hep_host (Philox stream, given seed) in the 10 cm
! LH2 cell, settings read from a raw file (20 float64: Ebeam, dEbeam, xwid,
! ywid, xoffset, yoffset, fr1, fr2, coulomb, e yptar min/max, e xptar
! min/max, e theta/phi/P, p theta/phi/P, Mh; then 3 int32: fr_pattern,
! using_eloss, using_coulomb).  Writes the 27 output columns (SoA float64)
! and the ok bytes.
!   usage: gen_events MAP_DIR N SEED SETTINGS_FILE OUT_FILE
program gen_events
  use iso_c_binding
  use simc_gpu
  implicit none
  character(len=512) :: map_dir, set_file, out_file, arg
  integer(c_int) :: rc
  integer(c_int64_t) :: n, seed
  real(c_double) :: par(20)
  integer(c_int32_t) :: ipar(3)
  real(c_double), allocatable, target :: o(:, :)
  integer(c_int8_t), allocatable, target :: ok(:)
  type(c_ptr) :: ctx
  type(simc_target) :: t
  type(simc_hep_settings) :: h
  type(simc_hep_out) :: hout
  type(simc_rng) :: rng
  integer :: u

  call get_command_argument(1, map_dir)
  call get_command_argument(2, arg); read (arg, *) n
  call get_command_argument(3, arg); read (arg, *) seed
  call get_command_argument(4, set_file)
  call get_command_argument(5, out_file)
  open (newunit=u, file=trim(set_file), access='stream', form='unformatted', status='old')
  read (u) par, ipar
  close (u)

  rc = simc_ctx_create(trim(map_dir)//c_null_char, ctx)
  if (rc /= SIMC_OK) then
    print *, 'simc_ctx_create failed: ', simc_error_message()
    stop 1
  end if

  ! the 2017 10 cm LH2 cell (X0 of hydrogen, init.f)
  t%A = 1; t%Z = 1; t%rho = 0.0723d0; t%X0 = 61.28d0; t%length = 10; t%angle = 0; t%zoffset = 0; t%can = 3
  h%Ebeam = par(1); h%dEbeam = par(2); h%xwid = par(3); h%ywid = par(4)
  h%xoffset = par(5); h%yoffset = par(6); h%fr1 = par(7); h%fr2 = par(8); h%coulomb = par(9)
  h%e_yptar = par(10:11); h%e_xptar = par(12:13)
  h%e_theta = par(14); h%e_phi = par(15); h%e_P = par(16)
  h%p_theta = par(17); h%p_phi = par(18); h%p_P = par(19); h%Mh = par(20)
  h%fr_pattern = ipar(1); h%using_eloss = ipar(2); h%using_coulomb = ipar(3)

  allocate(o(n, 27), ok(n))
  hout%x = c_loc(o(1, 1)); hout%y = c_loc(o(1, 2)); hout%z = c_loc(o(1, 3))
  hout%rasterx = c_loc(o(1, 4)); hout%rastery = c_loc(o(1, 5)); hout%Ein = c_loc(o(1, 6))
  hout%eloss = c_loc(o(1, 7)); hout%teff = c_loc(o(1, 8))
  hout%e_xptar = c_loc(o(1, 9)); hout%e_yptar = c_loc(o(1, 10)); hout%e_theta = c_loc(o(1, 11))
  hout%e_phi = c_loc(o(1, 12)); hout%e_E = c_loc(o(1, 13)); hout%e_delta = c_loc(o(1, 14))
  hout%p_xptar = c_loc(o(1, 15)); hout%p_yptar = c_loc(o(1, 16)); hout%p_theta = c_loc(o(1, 17))
  hout%p_phi = c_loc(o(1, 18)); hout%p_P = c_loc(o(1, 19)); hout%p_E = c_loc(o(1, 20))
  hout%p_delta = c_loc(o(1, 21)); hout%Q2 = c_loc(o(1, 22)); hout%nu = c_loc(o(1, 23))
  hout%q = c_loc(o(1, 24)); hout%jacobian = c_loc(o(1, 25)); hout%sigcc = c_loc(o(1, 26))
  hout%weight = c_loc(o(1, 27)); hout%ok = c_loc(ok(1))
  rng%kind = SIMC_RNG_PHILOX
  rng%seed = seed
  rng%event0 = 0
  rc = simc_gen_hep_host(ctx, t, h, hout, n, rng)
  if (rc /= SIMC_OK) then
    print *, 'simc_gen_hep_host failed: ', simc_error_message()
    stop 1
  end if
  open (newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
  write (u) o, ok
  close (u)
  print '(a, i0, a, es24.16)', 'generated ', n, ' events, weight sum ', sum(o(:, 27))
  call simc_ctx_destroy(ctx)
end program
