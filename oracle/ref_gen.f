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

c     ref_gen_one -- one event of the reference's generate (event.f) for
c     H(e,e'p) elastic without radiation, then complete_main for its weight
c     (recon = vertex).  par: Ebeam, dEbeam, xwid, ywid, xoffset, yoffset,
c     fr1, fr2, Coulomb_constant, gen e yptar min/max, xptar min/max,
c     spec e theta/phi/P, spec p theta/phi/P, Mh; ipar: fr_pattern,
c     using_Eloss, using_Coulomb.  out(1..27) in the order of the
c     generator's outputs; ok = generate's success.  The target
c     (/target_info/ A..can) is set by ref_set_target beforehand.
      subroutine ref_gen_one(par, ipar, out, ok)
      USE structureModule
      implicit none
      include 'simulate.inc'
      real*8 par(20), out(27)
      integer ipar(3), ok, i
      type(event_main):: main
      type(event):: vertex, orig, recon
      logical success
      Ebeam = par(1)
      dEbeam = par(2)
      gen%xwid = par(3)
      gen%ywid = par(4)
      targ%xoffset = par(5)
      targ%yoffset = par(6)
      targ%fr1 = par(7)
      targ%fr2 = par(8)
      targ%Coulomb_constant = par(9)
      gen%e%yptar%min = par(10)
      gen%e%yptar%max = par(11)
      gen%e%xptar%min = par(12)
      gen%e%xptar%max = par(13)
      spec%e%theta = par(14)
      spec%e%phi = par(15)
      spec%e%P = par(16)
      spec%p%theta = par(17)
      spec%p%phi = par(18)
      spec%p%P = par(19)
      Mh = par(20)
      Mh2 = Mh*Mh
      targ%fr_pattern = ipar(1)
      using_Eloss = ipar(2).ne.0
      using_Coulomb = ipar(3).ne.0
      using_rad = .false.
      doing_hyd_elast = .true.
      doing_deuterium = .false.
      doing_heavy = .false.
      doing_pion = .false.
      doing_kaon = .false.
      doing_delta = .false.
      doing_rho = .false.
      doing_semi = .false.
      doing_phsp = .false.
      do i = 1, 6
        debug(i) = .false.
      enddo
      call generate(main, vertex, orig, success)
      ok = 0
      if (success) ok = 1
      recon = vertex
      main%weight = 0
      if (success) call complete_main(.false., main, vertex, vertex,
     >                                recon, success)
      out(1) = main%target%x
      out(2) = main%target%y
      out(3) = main%target%z
      out(4) = main%target%rasterx
      out(5) = main%target%rastery
      out(6) = vertex%Ein
      out(7) = main%target%Eloss(1)
      out(8) = main%target%teff(1)
      out(9) = vertex%e%xptar
      out(10) = vertex%e%yptar
      out(11) = vertex%e%theta
      out(12) = vertex%e%phi
      out(13) = vertex%e%E
      out(14) = vertex%e%delta
      out(15) = vertex%p%xptar
      out(16) = vertex%p%yptar
      out(17) = vertex%p%theta
      out(18) = vertex%p%phi
      out(19) = vertex%p%P
      out(20) = vertex%p%E
      out(21) = vertex%p%delta
      out(22) = vertex%Q2
      out(23) = vertex%nu
      out(24) = vertex%q
      out(25) = main%jacobian
      out(26) = main%sigcc
      out(27) = main%weight
      return
      end
