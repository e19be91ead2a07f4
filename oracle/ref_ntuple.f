! ref_ntuple.f -- TEST INFRASTRUCTURE ONLY (not reference code).  Compiled
! against the reference's include files by oracle/Makefile into _ref, next to
! the reference's own NtupleInit.f and NtupleClose.f:
!   ref_ntuple_mode   sets the analysis-mode flags NtupleInit reads
!                     (simulate.inc), so that it writes that mode's header;
!   ref_ntuple_event  writes one event's values the way results_ntu_write
!                     writes its ntu array (one record per value,
!                     results_write.f:262-264).
	subroutine ref_ntuple_mode(mode, tgtfield)
	USE structureModule
	implicit none
	include 'simulate.inc'
	integer*4 mode, tgtfield
	doing_hyd_elast = mode.eq.1
	doing_deuterium = .false.
	doing_heavy = .false.
	doing_pion = mode.eq.2
	doing_kaon = .false.
	doing_delta = .false.
	doing_semi = mode.eq.3
	doing_rho = .false.
	doing_pizero = .false.
	using_tgt_field = tgtfield.ne.0
	end

	subroutine ref_ntuple_event(v)
	implicit none
	include 'hbook.inc'
	real*8 v(*)
	integer*4 i
	do i = 1, NtupleSize
	  write(NtupleIO) v(i)
	enddo
	end
