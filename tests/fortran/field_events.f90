! The target-field drop-ins called like simc.f does: trgInit, then per event
! track_from_tgt (simc.f:1430) and, with /track/ holding the focal-plane
! track, track_to_tgt (simc.f:1577-1583), linked against libsimc_dropin.so.
! Reads n events: from-target inputs (SoA float64 x, y, z, dx, dy, mom) and
! to-target inputs (x_fp, dx_fp, y_fp, dy_fp, delta, y, dx, dy, frx, fry,
! mom); writes x, y, z, dx, dy of the first and delta, y, dx, dy of the
! second (float64 SoA), then the two 4-byte ok flags per event.
!   usage: field_events MAP THETA_E THETA_P ARM SPECT MASS CTHETA STHETA N IN OUT   (SIMC_MAP_DIR set)
program field_events
  implicit none
  real*8 trk(80)
  common /track/ trk
  character(len=512) :: map, in_file, out_file, arg
  integer :: arm, spect, u, i
  integer*8 :: n
  real*8 :: theta_e, theta_p, mass, ctheta, stheta, zero
  real*8, allocatable :: a(:, :), b(:, :), oa(:, :), ob(:, :)
  integer*4, allocatable :: oka(:), okb(:)
  logical :: ok

  call get_command_argument(1, map)
  call get_command_argument(2, arg); read (arg, *) theta_e
  call get_command_argument(3, arg); read (arg, *) theta_p
  call get_command_argument(4, arg); read (arg, *) arm
  call get_command_argument(5, arg); read (arg, *) spect
  call get_command_argument(6, arg); read (arg, *) mass
  call get_command_argument(7, arg); read (arg, *) ctheta
  call get_command_argument(8, arg); read (arg, *) stheta
  call get_command_argument(9, arg); read (arg, *) n
  call get_command_argument(10, in_file)
  call get_command_argument(11, out_file)
  allocate(a(n, 6), b(n, 11), oa(n, 5), ob(n, 4), oka(n), okb(n))
  open (newunit=u, file=trim(in_file), access='stream', form='unformatted', status='old')
  read (u) a, b
  close (u)
  zero = 0
  call trginit(map, theta_e, zero, theta_p, zero)
  do i = 1, n
    oa(i, :) = a(i, 1:5)
    call track_from_tgt(oa(i, 1), oa(i, 2), oa(i, 3), oa(i, 4), oa(i, 5), a(i, 6), mass, spect, ok)
    oka(i) = 0
    if (ok) oka(i) = 1
  end do
  do i = 1, n
    trk(1) = b(i, 1)     ! xs
    trk(5) = b(i, 2)     ! dxdzs
    trk(2) = b(i, 3)     ! ys
    trk(6) = b(i, 4)     ! dydzs
    ob(i, :) = b(i, 5:8)
    ok = .true.
    call track_to_tgt(ob(i, 1), ob(i, 2), ob(i, 3), ob(i, 4), b(i, 9), b(i, 10), b(i, 11), mass, ctheta, &
                      stheta, spect, ok, arm)
    okb(i) = 0
    if (ok) okb(i) = 1
  end do
  call simc_dropin_close()
  open (newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
  write (u) oa, ob, oka, okb
  close (u)
end program field_events
