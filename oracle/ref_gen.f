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

c     ref_gen_rad_one -- one event of the reference's generate (event.f)
c     for H(e,e'p) elastic WITH radiation (generate_rad, radc.f), then
c     complete_main for its weight.  par, ipar as ref_gen_one; rpar:
c     dE_edge_test, Egamma_gen_max, edge%e%E min/max, edge%p%E min/max,
c     edge%Em min/max, VERTEXedge%Trec min/max, Egamma_tot_max; irpar:
c     rad_flag, extrad_flag, intcor_mode, use_expon, use_offshell_rad,
c     doing_tail(1..3).  The run constants of radc_init (init.f:576-651)
c     come from the reference's own radc_init; target_init's L1, L2 as
c     init.f:19-24.  out(1..27) as ref_gen_one, then out(28..44):
c     Egamma_used(1..3), ntup%radarm, main%gen_weight, orig%Ein,
c     orig%e%E, orig%e%delta, orig%p%E, orig%p%P, orig%p%delta,
c     main%target%Eloss(2), teff(2), Eloss(3), teff(3), hardcorfac,
c     etatzai.
      subroutine ref_gen_rad_one(par, ipar, rpar, irpar, out, ok)
      USE structureModule
      implicit none
      include 'simulate.inc'
      include 'radc.inc'
      real*8 par(20), rpar(11), out(44)
      integer ipar(3), irpar(8), ok, i
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
      using_rad = .true.
      doing_eep = .true.
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
      targ%L1 = log(184.15) - log(targ%Z)/3.0
      targ%L2 = log(1194.) - 2.*log(targ%Z)/3.0
      if (targ%Z.eq.1) then
        targ%L1 = 5.31
        targ%L2 = 6.144
      endif
      dE_edge_test = rpar(1)
      Egamma_gen_max = rpar(2)
      hardwired_rad = Egamma_gen_max.gt.0.01
      edge%e%E%min = rpar(3)
      edge%e%E%max = rpar(4)
      edge%p%E%min = rpar(5)
      edge%p%E%max = rpar(6)
      edge%Em%min = rpar(7)
      edge%Em%max = rpar(8)
      VERTEXedge%Trec%min = rpar(9)
      VERTEXedge%Trec%max = rpar(10)
      Egamma_tot_max = rpar(11)
      rad_flag = irpar(1)
      extrad_flag = irpar(2)
      intcor_mode = irpar(3)
      use_expon = irpar(4)
      use_offshell_rad = irpar(5).ne.0
      do i = 1, 3
        doing_tail(i) = irpar(5+i).ne.0
      enddo
      Egamma1_max = 0.0
      Egamma2_max = 0.0
      Egamma3_max = 0.0
      if (doing_tail(1)) Egamma1_max = Egamma_tot_max
      if (doing_tail(2)) Egamma2_max = Egamma_tot_max
      if (doing_tail(3)) Egamma3_max = Egamma_tot_max
      call radc_init
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
      do i = 1, 3
        out(27+i) = Egamma_used(i)
      enddo
      out(31) = ntup%radarm
      out(32) = main%gen_weight
      out(33) = orig%Ein
      out(34) = orig%e%E
      out(35) = orig%e%delta
      out(36) = orig%p%E
      out(37) = orig%p%P
      out(38) = orig%p%delta
      out(39) = main%target%Eloss(2)
      out(40) = main%target%teff(2)
      out(41) = main%target%Eloss(3)
      out(42) = main%target%teff(3)
      out(43) = hardcorfac
      out(44) = etatzai
      return
      end

c     ref_rad_setflags -- the brem switches as radc_init sets them
c     (init.f:637-648) for use_expon, for direct calls of brem/bremos.
      subroutine ref_rad_setflags(iexpon)
      implicit none
      include 'brem.inc'
      integer iexpon
      produce_output = .false.
      exponentiate = iexpon.eq.1
      include_hard = .true.
      calculate_spence = .true.
      return
      end
