c     ref_gen.f -- calls the reference's sigep (physics_proton.f) on a
c     vertex whose Q2, e%E, e%theta and Ein are given, and fofa_best_fit.
c     TEST INFRASTRUCTURE ONLY: compiled into oracle/_ref against the
c     reference tree's include files.
      real*8 function ref_sigep(q2, e, theta, ein)
      USE structureModule
      implicit none
      include 'simulate.inc'
      real*8 q2, e, theta, ein, sigep
      type(event):: v
      v%Q2 = q2
      v%e%E = e
      v%e%theta = theta
      v%Ein = ein
      ref_sigep = sigep(v)
      return
      end

      subroutine ref_fofa(qsquar, ge, gm)
      implicit none
      real*8 qsquar, ge, gm
      call fofa_best_fit(qsquar, ge, gm)
      return
      end
