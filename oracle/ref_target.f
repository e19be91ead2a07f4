c     ref_target.f -- sets the reference's target description (common
c     /target_info/ of target.inc) and the spectrometer numbers of the two
c     arms (electron_arm, hadron_arm of /gnrl/, simulate.inc:76) before the
c     harness calls trip_thru_target and target_musc of target.f.
c     TEST INFRASTRUCTURE ONLY: compiled into oracle/_ref against the
c     reference tree's include files.  X0_cm as init.f:43.
      subroutine ref_set_target(a, z, rho, x0, tlen, tang, zoff, ican,
     >                          iearm, iharm)
      USE structureModule
      implicit none
      include 'simulate.inc'
      real*8 a, z, rho, x0, tlen, tang, zoff
      integer ican, iearm, iharm
      targ%A = a
      targ%Z = z
      targ%rho = rho
      targ%X0 = x0
      targ%X0_cm = x0/rho
      targ%length = tlen
      targ%angle = tang
      targ%zoffset = zoff
      targ%can = ican
      electron_arm = iearm
      hadron_arm = iharm
      return
      end
