! Fortran interface (ISO_C_BINDING) of libsimc_gpu.so, the GPU spectrometer
! transport + reconstruction library declared in include/simc_gpu.h.
!
! A SIMC program that batches its events calls simc_arm_run_host (host
! arrays) or simc_arm_run (device arrays, e.g. from hipfort) in place of one
! mc_hms / mc_shms call per event (hms/mc_hms.f:1-4, shms/mc_shms.f:1-4);
! simc_counters_get returns the STOP counters in the order of the common
! blocks /spectrometer1/ and /spectrometer_shms/.  Every function returns 0
! or a negative SIMC_E_* code; simc_error_message() describes the last error.
module simc_gpu
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: SIMC_OK = 0
  integer(c_int), parameter :: SIMC_ARM_HMS = 1, SIMC_ARM_SOS = 2, SIMC_ARM_HRSR = 3, SIMC_ARM_HRSL = 4
  integer(c_int), parameter :: SIMC_ARM_SHMS = 5
  integer(c_int32_t), parameter :: SIMC_RNG_TAPE = 0, SIMC_RNG_PHILOX = 1
  integer(c_int32_t), parameter :: SIMC_FIT_FP64 = 0, SIMC_FIT_REF_F32 = 1
  integer(c_int), parameter :: SIMC_HMS_NCOUNTERS = 22, SIMC_SHMS_NCOUNTERS = 52, SIMC_SOS_NCOUNTERS = 18
  integer(c_int), parameter :: SIMC_HRS_NCOUNTERS = 21

  type, bind(C) :: simc_settings
    real(c_double)     :: p_spec = 0, m2 = 0, ctau = 0
    integer(c_int32_t) :: ms_flag = 1, wcs_flag = 1, decay_flag = 0, using_coll = 0
    integer(c_int32_t) :: fit_mode = SIMC_FIT_REF_F32, sieve = 0
    real(c_double)     :: wire_sigma = 0
  end type

  type, bind(C) :: simc_events_in
    type(c_ptr) :: dpp = c_null_ptr, x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: dxdz = c_null_ptr, dydz = c_null_ptr, fry = c_null_ptr
  end type

  type, bind(C) :: simc_events_out
    type(c_ptr) :: dpp = c_null_ptr, y = c_null_ptr, dxdz = c_null_ptr, dydz = c_null_ptr
    type(c_ptr) :: x_fp = c_null_ptr, dx_fp = c_null_ptr, y_fp = c_null_ptr, dy_fp = c_null_ptr
    type(c_ptr) :: pathlen = c_null_ptr, resmult = c_null_ptr, m2 = c_null_ptr
    type(c_ptr) :: ok = c_null_ptr, stop = c_null_ptr, used = c_null_ptr
  end type

  type, bind(C) :: simc_rng
    integer(c_int32_t) :: kind = SIMC_RNG_PHILOX, reserved = 0
    integer(c_int64_t) :: seed = 0, event0 = 0
    type(c_ptr)        :: tape = c_null_ptr
    integer(c_int64_t) :: stride = 0
  end type

  ! one kinematic setting of a batched run (simc_batch_create)
  type, bind(C) :: simc_setting
    integer(c_int32_t) :: arm = SIMC_ARM_HMS, reserved = 0
    type(simc_settings) :: s
    integer(c_int64_t) :: seed = 0, event0 = 0, offset = 0, count = 0
  end type

  ! target step (simc_target_prep): the target, one particle's preparation,
  ! the vertex arrays and the outputs (the arm's inputs plus loss, radiation
  ! lengths and scattering angles)
  integer(c_int32_t), parameter :: SIMC_TARGET_BEAM = 1, SIMC_TARGET_ELECTRON = 2, SIMC_TARGET_HADRON = 3

  type, bind(C) :: simc_target
    real(c_double)     :: A = 1, Z = 1, rho = 0.0723d0, X0 = 0, length = 10, angle = 0, zoffset = 0
    integer(c_int32_t) :: can = 3, reserved = 0
  end type

  type, bind(C) :: simc_prep_settings
    integer(c_int32_t) :: narm = SIMC_TARGET_HADRON, spectrometer = SIMC_ARM_HMS, using_eloss = 1, mc_smear = 1
    real(c_double)     :: p_spec = 0, th_spec = 0, phi = 0, mass = 0, coulomb = 0, beam_energy = 0
    real(c_double)     :: offset(5) = 0
  end type

  type, bind(C) :: simc_vertex_in
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr, E = c_null_ptr, p = c_null_ptr
    type(c_ptr) :: E_vertex = c_null_ptr, theta = c_null_ptr, delta = c_null_ptr, xptar = c_null_ptr
    type(c_ptr) :: yptar = c_null_ptr, dang_yp = c_null_ptr, dang_xp = c_null_ptr
  end type

  type, bind(C) :: simc_prep_out
    type(c_ptr) :: dpp = c_null_ptr, x = c_null_ptr, y = c_null_ptr, z = c_null_ptr, dxdz = c_null_ptr
    type(c_ptr) :: dydz = c_null_ptr, fry = c_null_ptr, eloss = c_null_ptr, teff = c_null_ptr
    type(c_ptr) :: dang1 = c_null_ptr, dang2 = c_null_ptr, used = c_null_ptr
  end type

  ! event generator (simc_gen_hep): H(e,e'p) elastic, radiation off
  type, bind(C) :: simc_hep_settings
    real(c_double)     :: Ebeam = 0, dEbeam = 0, xwid = 0, ywid = 0, xoffset = 0, yoffset = 0
    real(c_double)     :: fr1 = 0, fr2 = 0, coulomb = 0
    real(c_double)     :: e_yptar(2) = 0, e_xptar(2) = 0
    real(c_double)     :: e_theta = 0, e_phi = 0, e_P = 0, p_theta = 0, p_phi = 0, p_P = 0
    real(c_double)     :: Mh = 938.27231d0
    integer(c_int32_t) :: fr_pattern = 0, using_eloss = 1, using_coulomb = 0, reserved = 0
  end type

  type, bind(C) :: simc_hep_out
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr, rasterx = c_null_ptr, rastery = c_null_ptr
    type(c_ptr) :: Ein = c_null_ptr, eloss = c_null_ptr, teff = c_null_ptr
    type(c_ptr) :: e_xptar = c_null_ptr, e_yptar = c_null_ptr, e_theta = c_null_ptr, e_phi = c_null_ptr
    type(c_ptr) :: e_E = c_null_ptr, e_delta = c_null_ptr
    type(c_ptr) :: p_xptar = c_null_ptr, p_yptar = c_null_ptr, p_theta = c_null_ptr, p_phi = c_null_ptr
    type(c_ptr) :: p_P = c_null_ptr, p_E = c_null_ptr, p_delta = c_null_ptr
    type(c_ptr) :: Q2 = c_null_ptr, nu = c_null_ptr, q = c_null_ptr, jacobian = c_null_ptr
    type(c_ptr) :: sigcc = c_null_ptr, weight = c_null_ptr, ok = c_null_ptr, used = c_null_ptr
  end type

  ! generate with using_rad (radc.f, brem.f): simc_rad_settings, simc_rad_out
  type, bind(C) :: simc_rad_settings
    integer(c_int32_t) :: rad_flag = 0, extrad_flag = 0, intcor_mode = 1, use_expon = 0
    integer(c_int32_t) :: use_offshell_rad = 0, doing_tail(3) = 1, e_arm = 5, p_arm = 1
    real(c_double)     :: dE_edge_test = 0, Egamma_gen_max = 0
    real(c_double)     :: edge_e_E(2) = 0, edge_p_E(2) = 0, edge_Em(2) = 0, vertex_Trec(2) = 0
    real(c_double)     :: Egamma_tot_max = 0, etatzai = 0
  end type

  type, bind(C) :: simc_rad_out
    type(c_ptr) :: Egamma1 = c_null_ptr, Egamma2 = c_null_ptr, Egamma3 = c_null_ptr, gen_weight = c_null_ptr
    type(c_ptr) :: orig_Ein = c_null_ptr, orig_e_E = c_null_ptr, orig_e_delta = c_null_ptr
    type(c_ptr) :: orig_p_E = c_null_ptr, orig_p_P = c_null_ptr, orig_p_delta = c_null_ptr
    type(c_ptr) :: eloss2 = c_null_ptr, teff2 = c_null_ptr, eloss3 = c_null_ptr, teff3 = c_null_ptr
    type(c_ptr) :: hardcorfac = c_null_ptr, ntail = c_null_ptr
  end type

  interface
    integer(c_int) function simc_gen_hep(ctx, t, h, hout, n, rng, stream) bind(C, name="simc_gen_hep")
      import :: c_int, c_int64_t, c_ptr, simc_target, simc_hep_settings, simc_hep_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_target), intent(in) :: t
      type(simc_hep_settings), intent(in) :: h
      type(simc_hep_out), intent(in) :: hout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_gen_hep_rad(ctx, t, h, r, hout, rout, n, rng, stream) &
        bind(C, name="simc_gen_hep_rad")
      import :: c_int, c_int64_t, c_ptr, simc_target, simc_hep_settings, simc_rad_settings, simc_hep_out, &
                simc_rad_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_target), intent(in) :: t
      type(simc_hep_settings), intent(in) :: h
      type(simc_rad_settings), intent(in) :: r
      type(simc_hep_out), intent(in) :: hout
      type(simc_rad_out), intent(in) :: rout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_gen_hep_host(ctx, t, h, hout, n, rng) bind(C, name="simc_gen_hep_host")
      import :: c_int, c_int64_t, c_ptr, simc_target, simc_hep_settings, simc_hep_out, simc_rng
      type(c_ptr), value :: ctx
      type(simc_target), intent(in) :: t
      type(simc_hep_settings), intent(in) :: h
      type(simc_hep_out), intent(in) :: hout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_target_prep(ctx, t, s, vin, pout, n, rng, stream) bind(C, name="simc_target_prep")
      import :: c_int, c_int64_t, c_ptr, simc_target, simc_prep_settings, simc_vertex_in, simc_prep_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_target), intent(in) :: t
      type(simc_prep_settings), intent(in) :: s
      type(simc_vertex_in), intent(in) :: vin
      type(simc_prep_out), intent(in) :: pout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_target_prep_host(ctx, t, s, vin, pout, n, rng) &
        bind(C, name="simc_target_prep_host")
      import :: c_int, c_int64_t, c_ptr, simc_target, simc_prep_settings, simc_vertex_in, simc_prep_out, simc_rng
      type(c_ptr), value :: ctx
      type(simc_target), intent(in) :: t
      type(simc_prep_settings), intent(in) :: s
      type(simc_vertex_in), intent(in) :: vin
      type(simc_prep_out), intent(in) :: pout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    ! target-field tracking (trg_track.f): map = c_null_char-terminated file name,
    ! '0' for no field, blanks for the uniform test field (as trgInit)
    integer(c_int) function simc_field_create(map, theta_e, phi_e, theta_p, phi_p, f) &
        bind(C, name="simc_field_create")
      import :: c_int, c_char, c_double, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: map
      real(c_double), value :: theta_e, phi_e, theta_p, phi_p
      type(c_ptr), intent(out) :: f
    end function

    subroutine simc_field_destroy(f) bind(C, name="simc_field_destroy")
      import :: c_ptr
      type(c_ptr), value :: f
    end subroutine

    integer(c_int) function simc_track_from_tgt(f, spect, n, x, y, z, dx, dy, mom, mass, ok, stream) &
        bind(C, name="simc_track_from_tgt")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: f, x, y, z, dx, dy, mom, ok, stream
      integer(c_int), value :: spect
      integer(c_int64_t), value :: n
      real(c_double), value :: mass
    end function

    integer(c_int) function simc_track_to_tgt(ctx, f, arm, spect, n, x_fp, dx_fp, y_fp, dy_fp, delta, y, dx, dy, &
        frx, fry, mom, mass, ctheta, stheta, ok, stream) bind(C, name="simc_track_to_tgt")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, f, x_fp, dx_fp, y_fp, dy_fp, delta, y, dx, dy, frx, fry, mom, ok, stream
      integer(c_int), value :: arm, spect
      integer(c_int64_t), value :: n
      real(c_double), value :: mass, ctheta, stheta
    end function

    integer(c_int) function simc_track_from_tgt_host(f, spect, n, x, y, z, dx, dy, mom, mass, ok) &
        bind(C, name="simc_track_from_tgt_host")
      import :: c_int, c_int64_t, c_double, c_int8_t, c_ptr
      type(c_ptr), value :: f
      integer(c_int), value :: spect
      integer(c_int64_t), value :: n
      real(c_double), dimension(*), intent(inout) :: x, y, z, dx, dy
      real(c_double), dimension(*), intent(in) :: mom
      real(c_double), value :: mass
      integer(c_int8_t), dimension(*), intent(out) :: ok
    end function

    integer(c_int) function simc_track_to_tgt_host(ctx, f, arm, spect, n, x_fp, dx_fp, y_fp, dy_fp, delta, y, dx, &
        dy, frx, fry, mom, mass, ctheta, stheta, ok) bind(C, name="simc_track_to_tgt_host")
      import :: c_int, c_int64_t, c_double, c_int8_t, c_ptr
      type(c_ptr), value :: ctx, f
      integer(c_int), value :: arm, spect
      integer(c_int64_t), value :: n
      real(c_double), dimension(*), intent(in) :: x_fp, dx_fp, y_fp, dy_fp, frx, fry, mom
      real(c_double), dimension(*), intent(inout) :: delta, y, dx, dy
      real(c_double), value :: mass, ctheta, stheta
      integer(c_int8_t), dimension(*), intent(inout) :: ok
    end function

    ! output formats: the .bin ntuple (tags: array of c_loc of null-terminated
    ! names) and the STOP-counter report
    integer(c_int) function simc_ntuple_open(path, ncols, tags, nt) bind(C, name="simc_ntuple_open")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: ncols
      type(c_ptr), dimension(*), intent(in) :: tags
      type(c_ptr), intent(out) :: nt
    end function

    integer(c_int) function simc_ntuple_write(nt, rows, nrows) bind(C, name="simc_ntuple_write")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: nt
      real(c_double), dimension(*), intent(in) :: rows
      integer(c_int64_t), value :: nrows
    end function

    integer(c_int) function simc_ntuple_write_device(nt, cols, sel, n, written, stream) &
        bind(C, name="simc_ntuple_write_device")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: nt, sel, stream
      type(c_ptr), dimension(*), intent(in) :: cols
      integer(c_int64_t), value :: n
      integer(c_int64_t), intent(out) :: written
    end function

    integer(c_int) function simc_ntuple_close(nt) bind(C, name="simc_ntuple_close")
      import :: c_int, c_ptr
      type(c_ptr), value :: nt
    end function

    integer(c_int) function simc_stop_report(ctx, path, electron_arm, hadron_arm) bind(C, name="simc_stop_report")
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: electron_arm, hadron_arm
    end function

    integer(c_int) function simc_ctx_create(map_dir, ctx) bind(C, name="simc_ctx_create")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: map_dir
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function simc_ctx_create_files(hms_forward, hms_recon, shms_forward, shms_recon, ctx) &
        bind(C, name="simc_ctx_create_files")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: hms_forward, hms_recon, shms_forward, shms_recon
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function simc_ctx_load_arm(ctx, arm, forward, recon) bind(C, name="simc_ctx_load_arm")
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm
      character(kind=c_char), dimension(*), intent(in) :: forward, recon
    end function

    subroutine simc_ctx_destroy(ctx) bind(C, name="simc_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function simc_arm_run(ctx, arm, s, evin, evout, n, rng, stream) bind(C, name="simc_arm_run")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: arm
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    ! allocation-free variant: at most SIMC_NOALLOC_MAX_EVENTS (524288) events per call
    integer(c_int) function simc_arm_run_noalloc(ctx, arm, s, evin, evout, n, rng, stream) &
        bind(C, name="simc_arm_run_noalloc")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: arm
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_hms_run(ctx, s, evin, evout, n, rng, stream) bind(C, name="simc_hms_run")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_shms_run(ctx, s, evin, evout, n, rng, stream) bind(C, name="simc_shms_run")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_sos_run(ctx, s, evin, evout, n, rng, stream) bind(C, name="simc_sos_run")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_hrsr_run(ctx, s, evin, evout, n, rng, stream) bind(C, name="simc_hrsr_run")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_hrsl_run(ctx, s, evin, evout, n, rng, stream) bind(C, name="simc_hrsl_run")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_arm_run_host(ctx, arm, s, evin, evout, n, rng) bind(C, name="simc_arm_run_host")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
    end function

    integer(c_int) function simc_recon_run(ctx, arm, n, x_fp, dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, &
        dydz_tar, y_tar, stream) bind(C, name="simc_recon_run")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, x_fp, dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, dydz_tar, y_tar, stream
      integer(c_int), value :: arm
      integer(c_int64_t), value :: n
    end function

    integer(c_int) function simc_recon_run_host(ctx, arm, n, x_fp, dx_fp, y_fp, dy_fp, fry, delta, dxdz_tar, &
        dydz_tar, y_tar) bind(C, name="simc_recon_run_host")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x_fp(*), dx_fp(*), y_fp(*), dy_fp(*), fry(*)
      real(c_double), intent(out) :: delta(*), dxdz_tar(*), dydz_tar(*), y_tar(*)
    end function

    integer(c_int) function simc_arm_run_timed(ctx, arm, s, evin, evout, n, rng, stream, stage_events) &
        bind(C, name="simc_arm_run_timed")
      import :: c_int, c_int64_t, c_ptr, simc_settings, simc_events_in, simc_events_out, simc_rng
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: arm
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
      type(c_ptr), intent(in) :: stage_events(5)
    end function

    integer(c_int) function simc_map_class_info(ctx, arm, cls, driftdist, adrift, n_terms) &
        bind(C, name="simc_map_class_info")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm, cls
      real(c_double), intent(out) :: driftdist
      integer(c_int32_t), intent(out) :: adrift, n_terms
    end function

    integer(c_int64_t) function simc_poly_flops_full(ctx, arm) bind(C, name="simc_poly_flops_full")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm
    end function

    integer(c_int) function simc_stop_flops(ctx, arm, flops, nflops) bind(C, name="simc_stop_flops")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm, nflops
      integer(c_int64_t), intent(out) :: flops(*)
    end function

    integer(c_int) function simc_stop_flops_stage(ctx, arm, stage, flops, nflops) &
        bind(C, name="simc_stop_flops_stage")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm, stage, nflops
      integer(c_int64_t), intent(out) :: flops(*)
    end function

    integer(c_int) function simc_stop_flops_hut(ctx, arm, ms_flag, wcs_flag, flops, nflops) &
        bind(C, name="simc_stop_flops_hut")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm, ms_flag, wcs_flag, nflops
      integer(c_int64_t), intent(out) :: flops(*)
    end function

    integer(c_int) function simc_adjoint_run(ctx, arm, s, evin, mask, n, rng, recon_coef, offsets, grad_coef, &
        out9, stream) bind(C, name="simc_adjoint_run")
      import :: c_int, c_int64_t, c_ptr, c_double, simc_settings, simc_events_in, simc_rng
      type(c_ptr), value :: ctx, mask, recon_coef, grad_coef, out9, stream
      integer(c_int), value :: arm
      type(simc_settings), intent(in) :: s
      type(simc_events_in), intent(in) :: evin
      integer(c_int64_t), value :: n
      type(simc_rng), intent(in) :: rng
      real(c_double), intent(in) :: offsets(6)
    end function

    integer(c_int) function simc_recon_coefficients(ctx, arm, coef, n_terms) bind(C, name="simc_recon_coefficients")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, coef
      integer(c_int), value :: arm
      integer(c_int32_t), intent(out) :: n_terms
    end function

    integer(c_int) function simc_batch_create(ctx, settings, n_settings, b) bind(C, name="simc_batch_create")
      import :: c_int, c_ptr, simc_setting
      type(c_ptr), value :: ctx
      type(simc_setting), intent(in) :: settings(*)
      integer(c_int), value :: n_settings
      type(c_ptr), intent(out) :: b
    end function

    integer(c_int) function simc_batch_run(b, evin, evout, stream) bind(C, name="simc_batch_run")
      import :: c_int, c_ptr, simc_events_in, simc_events_out
      type(c_ptr), value :: b, stream
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
    end function

    integer(c_int) function simc_batch_counters(b, setting, c, nc) bind(C, name="simc_batch_counters")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: setting, nc
      integer(c_int64_t), intent(out) :: c(*)
    end function

    integer(c_int) function simc_batch_set_chunk(b, chunk) bind(C, name="simc_batch_set_chunk")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int64_t), value :: chunk
    end function

    integer(c_int) function simc_batch_tune(b, evin, evout, stream, best_chunk, best_ms) &
        bind(C, name="simc_batch_tune")
      import :: c_int, c_int64_t, c_double, c_ptr, simc_events_in, simc_events_out
      type(c_ptr), value :: b, stream
      type(simc_events_in), intent(in) :: evin
      type(simc_events_out), intent(in) :: evout
      integer(c_int64_t), intent(out) :: best_chunk
      real(c_double), intent(out) :: best_ms
    end function

    subroutine simc_batch_destroy(b) bind(C, name="simc_batch_destroy")
      import :: c_ptr
      type(c_ptr), value :: b
    end subroutine

    integer(c_int) function simc_batch_passes(b, n_passes, n_lane_key) bind(C, name="simc_batch_passes")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), intent(out) :: n_passes, n_lane_key
    end function

    integer(c_int) function simc_counters_get(ctx, arm, c, nc) bind(C, name="simc_counters_get")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm, nc
      integer(c_int64_t), intent(out) :: c(*)
    end function

    integer(c_int) function simc_counters_reset(ctx, arm) bind(C, name="simc_counters_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: arm
    end function

    type(c_ptr) function simc_last_error() bind(C, name="simc_last_error")
      import :: c_ptr
    end function

    type(c_ptr) function simc_version() bind(C, name="simc_version")
      import :: c_ptr
    end function

    integer(c_size_t) function c_strlen(p) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: p
    end function
  end interface

contains

  ! The message of the last failed call on this thread.
  function simc_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = simc_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

end module simc_gpu
