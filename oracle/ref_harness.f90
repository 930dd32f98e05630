! Parity-checker harness around the reference CodeNS routines.
!
! TEST INFRASTRUCTURE ONLY.  This file is compiled together with a handful of
! the reference's own Fortran sources (read in place under the reference tree
! by oracle/Makefile) into oracle/_ref/libcodens_ref.so.  Only the tests, the
! smoke check and the CPU-baseline leg of bench.py load that library, and only
! as a checker or a timed CPU baseline: the product path never calls it.
!
! Each entry point sets the reference's module globals (the extents of the
! blocks, the gas constants, the scheme switches) from plain C arguments and
! then calls the reference routine unchanged, with the arrays passed by
! sequence association exactly as atsch_num.f90 passes them.
module codens_ref_harness
  use iso_c_binding
  use para_var
  use para_fige
  use maillage
  use proprieteflu
  use schemanum
  use definition
  use modeleturb
  implicit none
  character(len=7) :: heq = 'ns3d   '
contains

  ! --------------------------------------------------------------------
  ! Global layout: nb blocks, storage ip11 cells, largest block ip00.
  ! ext(16*nb): per block npc, id1,id2,jd1,jd2,kd1,kd2, ii1,ii2,jj1,jj2,
  ! kk1,kk2, nnn, (unused), (unused).   (crdms.f90:126-153)
  ! --------------------------------------------------------------------
  subroutine ref_set_layout(nb, nip11, nip00, ext) bind(C, name="ref_set_layout")
    integer(c_int), value :: nb, nip11, nip00
    integer(c_int) :: ext(16, *)
    integer :: l
    if (allocated(npc)) then
       deallocate(npc, npn, npfb, nnn, nnc, nnfb, id1, id2, jd1, jd2, kd1, kd2, &
            ii1, ii2, jj1, jj2, kk1, kk2)
    endif
    allocate(npc(nb), npn(nb), npfb(nb), nnn(nb), nnc(nb), nnfb(nb))
    allocate(id1(nb), id2(nb), jd1(nb), jd2(nb), kd1(nb), kd2(nb))
    allocate(ii1(nb), ii2(nb), jj1(nb), jj2(nb), kk1(nb), kk2(nb))
    do l = 1, nb
       npc(l) = ext(1, l)
       npn(l) = ext(1, l)
       npfb(l) = nind*ext(1, l)
       id1(l) = ext(2, l);  id2(l) = ext(3, l)
       jd1(l) = ext(4, l);  jd2(l) = ext(5, l)
       kd1(l) = ext(6, l);  kd2(l) = ext(7, l)
       ii1(l) = ext(8, l);  ii2(l) = ext(9, l)
       jj1(l) = ext(10, l); jj2(l) = ext(11, l)
       kk1(l) = ext(12, l); kk2(l) = ext(13, l)
       nnn(l) = ext(14, l)
       nnc(l) = ext(14, l)
       nnfb(l) = nind*ext(14, l)
    enddo
    lzx = nb
    lz = nb
    ip00 = nip00
    ip11 = nip11
    ip12 = nip11
    ip13 = nip11
    ip21 = nip11
    ip31 = 3*nip11
    ip42 = nip00
    ip60 = 7
  end subroutine ref_set_layout

  subroutine ref_set_phys(g, pinf, prl, prtl, rey, tref, iflul) bind(C, name="ref_set_phys")
    real(c_double), value :: g, pinf, prl, prtl, rey, tref
    integer(c_int), value :: iflul
    gam = g
    gam1 = gam - 1.
    gam2 = .5*gam1
    gam3 = 1./gam
    gam4 = 1./gam1
    gam5 = gam*gam1
    pinfl = pinf
    pr = prl
    prt = prtl
    reynz = rey
    tnz = tref
    iflu = iflul
    cp = 1./(gam - 1.)
    cv = cp/gam
  end subroutine ref_set_phys

  subroutine ref_set_scheme(imuscl, xkl, iliml, kfmgl, ins) bind(C, name="ref_set_scheme")
    integer(c_int), value :: imuscl, iliml, kfmgl, ins
    real(c_double), value :: xkl
    muscl = imuscl
    xk = xkl
    ilim = iliml
    kfmg = kfmgl
    if (ins .ne. 0) then
       heq = 'ns3d   '
    else
       heq = 'eu3d   '
    endif
  end subroutine ref_set_scheme

  ! zpres.f90:4
  subroutine ref_zpres(l, v, ps, temp, cson) bind(C, name="ref_zpres")
    use mod_zpres
    integer(c_int), value :: l
    real(c_double) :: v(*), ps(*), temp(*), cson(*)
    call zpres(l, v, ps, temp, cson)
  end subroutine ref_zpres

  ! zvismo.f90:4
  subroutine ref_zvismo(l, mu, v, temp) bind(C, name="ref_zvismo")
    use mod_zvismo
    integer(c_int), value :: l
    real(c_double) :: mu(*), v(*), temp(*)
    call zvismo(l, mu, v, temp)
  end subroutine ref_zvismo

  ! chronos.f90:4 (local time step of one block)
  subroutine ref_chronos(l, etal, mu, mut, v, dt, sn, lgsnlt, vol, t1, t2, t3, t4, cson) &
       bind(C, name="ref_chronos")
    use mod_chronos
    integer(c_int), value :: l, lgsnlt
    real(c_double), value :: etal
    real(c_double) :: mu(*), mut(*), v(*), dt(*), sn(*), vol(*), t1(*), t2(*), t3(*), t4(*), cson(*)
    call chronos(l, etal, mu, mut, v, dt, heq, sn, lgsnlt, vol, t1, t2, t3, t4, cson)
  end subroutine ref_chronos

  ! sch_roe_euler.f90:4
  subroutine ref_sch_roe_euler(l, ityprk, u, v, ff, sn, lgsnlt, t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps) &
       bind(C, name="ref_sch_roe_euler")
    use mod_sch_roe_euler
    integer(c_int), value :: l, ityprk, lgsnlt
    real(c_double) :: u(*), v(*), ff(*), sn(*), ps(*)
    real(c_double) :: t1(*), t2(*), t3(*), t4(*), t5(*), t6(*), t7(*), t8(*), t9(*), t10(*)
    call sch_roe_euler(l, ityprk, u, v, ff, heq, sn, lgsnlt, t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps)
  end subroutine ref_sch_roe_euler

  ! sch_roe.f90:4
  subroutine ref_sch_roe(l, ityprk, u, v, ff, txx, txy, txz, tyy, tyz, tzz, qx, qy, qz, &
       sn, lgsnlt, t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps) bind(C, name="ref_sch_roe")
    use mod_sch_roe
    integer(c_int), value :: l, ityprk, lgsnlt
    real(c_double) :: u(*), v(*), ff(*), sn(*), ps(*)
    real(c_double) :: txx(*), txy(*), txz(*), tyy(*), tyz(*), tzz(*), qx(*), qy(*), qz(*)
    real(c_double) :: t1(*), t2(*), t3(*), t4(*), t5(*), t6(*), t7(*), t8(*), t9(*), t10(*)
    call sch_roe(l, ityprk, u, v, ff, txx, txy, txz, tyy, tyz, tzz, qx, qy, qz, heq, sn, lgsnlt, &
         t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps)
  end subroutine ref_sch_roe

  ! zgrad2.f90:4 (cmui*/cmuj*/cmuk* weights are 1 outside the weighted schemes, inivec.f90:107-115)
  subroutine ref_zgrad2(l, sn, lgsnlt, vol, s, temp, g, cmu) bind(C, name="ref_zgrad2")
    use mod_zgrad2
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: sn(*), vol(*), s(*), temp(*), g(ip00, 12), cmu(*)
    call zgrad2(l, heq, sn, lgsnlt, vol, s, temp, &
         g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         g(1,10), g(1,11), g(1,12), cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_zgrad2

  ! zfluto.f90:4
  subroutine ref_zfluto(l, mu, mut, tq, g) bind(C, name="ref_zfluto")
    use mod_zfluto
    integer(c_int), value :: l
    real(c_double) :: mu(*), mut(*), tq(ip11, 9), g(ip00, 12)
    call zfluto(l, mu, mut, tq(1,1), tq(1,2), tq(1,3), tq(1,4), tq(1,5), tq(1,6), &
         tq(1,7), tq(1,8), tq(1,9), &
         g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         g(1,10), g(1,11), g(1,12))
  end subroutine ref_zfluto

  ! sch_expli.f90:4
  subroutine ref_sch_expli(l, u, v, dt, vol) bind(C, name="ref_sch_expli")
    use mod_sch_expli
    integer(c_int), value :: l
    real(c_double) :: u(*), v(*), dt(*), vol(*)
    call sch_expli(l, u, v, dt, vol)
  end subroutine ref_sch_expli

  ! implimf_3d.f90:4
  subroutine ref_implimf_3d(l, u, dt, v, d, ff, mu, mut, lmx, sn, lgsnlt, vol, dtpas, ityprk, &
       t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps, cson) bind(C, name="ref_implimf_3d")
    use mod_implimf_3d
    integer(c_int), value :: l, lmx, lgsnlt, ityprk
    real(c_double), value :: dtpas
    real(c_double) :: u(*), dt(*), v(*), d(*), ff(*), mu(*), mut(*), sn(*), vol(*), ps(*), cson(*)
    real(c_double) :: t1(*), t2(*), t3(*), t4(*), t5(*), t6(*), t7(*), t8(*), t9(*), t10(*)
    call implimf_3d(l, u, dt, v, d, ff, mu, mut, heq, lmx, sn, lgsnlt, vol, dtpas, ityprk, &
         t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps, cson)
  end subroutine ref_implimf_3d

  ! schemanum%dt1min (read by implimf_eu.f90:82 when kfmg = 3)
  subroutine ref_set_dt1min(x) bind(C, name="ref_set_dt1min")
    real(c_double), value :: x
    dt1min = x
  end subroutine ref_set_dt1min

  ! implimf_eu.f90:6
  subroutine ref_implimf_eu(l, u, dt, v, d, ff, lmx, sn, lgsnlt, vol, dtpas, ityprk, &
       t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps, cson) bind(C, name="ref_implimf_eu")
    use mod_implimf_eu
    integer(c_int), value :: l, lmx, lgsnlt, ityprk
    real(c_double), value :: dtpas
    real(c_double) :: u(*), dt(*), v(*), d(*), ff(*), sn(*), vol(*), ps(*), cson(*)
    real(c_double) :: t1(*), t2(*), t3(*), t4(*), t5(*), t6(*), t7(*), t8(*), t9(*), t10(*)
    call implimf_eu(l, u, dt, v, d, ff, heq, lmx, sn, lgsnlt, vol, dtpas, ityprk, &
         t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, ps, cson)
  end subroutine ref_implimf_eu

  ! norm.f90:4 via snorm.f90:4 (face area vectors of one block)
  subroutine ref_snorm(l, x, y, z, sn, lgsnlt) bind(C, name="ref_snorm")
    use mod_snorm
    use chainecarac
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: x(*), y(*), z(*), sn(*)
    equat = heq
    call snorm(l, x, y, z, sn, lgsnlt)
  end subroutine ref_snorm

  ! svol.f90:4
  subroutine ref_svol(l, x, y, z, sn, lgsnlt, vol, t1, t2, t3) bind(C, name="ref_svol")
    use mod_svol
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: x(*), y(*), z(*), sn(*), vol(*), t1(*), t2(*), t3(*)
    call svol(l, x, y, z, sn, lgsnlt, vol, t1, t2, t3)
  end subroutine ref_svol

  ! extmhg.f90:4
  subroutine ref_extmhg(l, x, y, z, ex1, ex2) bind(C, name="ref_extmhg")
    use mod_extmhg
    integer(c_int), value :: l
    real(c_double), value :: ex1, ex2
    real(c_double) :: x(*), y(*), z(*)
    call extmhg(l, x, y, z, ex1, ex2)
  end subroutine ref_extmhg

  ! dissip_jameson.f90:4
  subroutine ref_dissip_jameson(l, v, d, sn, lgsnlt, snu, ps, cson, ki2l, ki4l) &
       bind(C, name="ref_dissip_jameson")
    use mod_dissip_jameson
    integer(c_int), value :: l, lgsnlt
    real(c_double), value :: ki2l, ki4l
    real(c_double) :: v(*), d(*), sn(*), snu(*), ps(*), cson(*)
    if (.not. allocated(ki2)) allocate(ki2(size(npc)), ki4(size(npc)))
    ki2(l) = ki2l
    ki4(l) = ki4l
    call dissip_jameson(l, v, d, heq, sn, lgsnlt, snu, ps, cson)
  end subroutine ref_dissip_jameson

  ! met_roe2o.f90:4
  subroutine ref_met_roe2o(l, t, d, sn, lgsnlt, vol, del6, del7) bind(C, name="ref_met_roe2o")
    use mod_met_roe2o
    use chainecarac
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: t(*), d(*), sn(*), vol(*), del6(*), del7(*)
    equat = heq
    call met_roe2o(l, t, d, sn, lgsnlt, vol, del6, del7)
  end subroutine ref_met_roe2o

  ! met_roe.f90:4
  subroutine ref_met_roe(l, t, d, sn, lgsnlt, vol) bind(C, name="ref_met_roe")
    use mod_met_roe
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: t(*), d(*), sn(*), vol(*)
    call met_roe(l, t, d, heq, sn, lgsnlt, vol)
  end subroutine ref_met_roe

  ! modeleturb constants (atlecdon.f90:112-400); tc(16): cb1 cb2 sigma kappa
  ! cw1 cw2 cw3 cv1 ct3 ct4 betas beta sigmk sigmw alfak alfae
  subroutine ref_set_turb(tc) bind(C, name="ref_set_turb")
    real(c_double) :: tc(16)
    cb1 = tc(1); cb2 = tc(2); sigma = tc(3); kappa = tc(4); cw1 = tc(5); cw2 = tc(6)
    cw3 = tc(7); cv1 = tc(8); ct3 = tc(9); ct4 = tc(10); betas = tc(11); beta = tc(12)
    sigmk = tc(13); sigmw = tc(14); alfak = tc(15); alfae = tc(16)
    kparoi = 0
    lparoi = 0
    ischema = 7
  end subroutine ref_set_turb

  ! teq_gradv.f90:4 (grad u, v, w; clamps rho)
  subroutine ref_teq_gradv(l, sn, vol, v, s, g, cmu) bind(C, name="ref_teq_gradv")
    use mod_teq_gradv
    use chainecarac
    integer(c_int), value :: l
    real(c_double) :: sn(*), vol(*), v(*), s(*), g(ip00, 9), cmu(*)
    equat = heq
    call teq_gradv(l, sn, vol, v, s, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_teq_gradv

  ! met_prod.f90:4
  subroutine ref_met_prod(l, mut, v, g, tprod) bind(C, name="ref_met_prod")
    use mod_met_prod
    integer(c_int), value :: l
    real(c_double) :: mut(*), v(*), g(ip00, 9), tprod(*)
    call met_prod(l, 1, mut, v, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), tprod)
  end subroutine ref_met_prod

  ! met_gradtr.f90:4
  subroutine ref_met_gradtr(l, sn, vol, v, mu, mut, tt, gx, gy, gz, fd, cmu) bind(C, name="ref_met_gradtr")
    use mod_met_gradtr
    use chainecarac
    integer(c_int), value :: l
    real(c_double) :: sn(*), vol(*), v(*), mu(*), mut(*), tt(*), gx(*), gy(*), gz(*), fd(ip11, 6), cmu(*)
    equat = heq
    call met_gradtr(l, sn, vol, v, mu, mut, tt, gx, gy, gz, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), &
         fd(1,6), cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_gradtr

  ! met_smsa.f90:4 (calls met_difsa.f90:4)
  subroutine ref_met_smsa(l, sn, vol, v, mu, dist, g, fd, cfke, tt, gx, gy, gz, sdif, ts6, ts7, cmu) &
       bind(C, name="ref_met_smsa")
    use mod_met_smsa
    use chainecarac
    integer(c_int), value :: l
    real(c_double) :: sn(*), vol(*), v(*), mu(*), dist(*), g(ip00, 9), fd(ip11, 6), cfke(*), tt(*), gx(*), gy(*)
    real(c_double) :: gz(*), sdif(*), ts6(*), ts7(*), cmu(*)
    equat = heq
    call met_smsa(l, sn, vol, v, mu, dist, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), &
         g(1,9), fd(1,1), fd(1,2), fd(1,3), cfke, tt, gx, gy, gz, sdif, ts6, ts7, cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_smsa

  ! met_smko.f90:4
  subroutine ref_met_smko(l, v, fd, tprod, cfke, ts6, ts7, cson) bind(C, name="ref_met_smko")
    use mod_met_smko
    integer(c_int), value :: l
    real(c_double) :: v(*), fd(ip11, 6), tprod(*), cfke(*), ts6(*), ts7(*), cson(*)
    call met_smko(l, v, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), tprod, cfke, ts6, ts7, cson)
  end subroutine ref_met_smko

  ! met_fludcsa.f90:4, met_fludko.f90:4, met_fludc.f90:4 by kfludis
  subroutine ref_met_flud(l, kfl, v, mu, mut, fd, tt, gx, gy, gz) bind(C, name="ref_met_flud")
    use mod_met_fludcsa
    use mod_met_fludko
    use mod_met_fludc
    use chainecarac
    integer(c_int), value :: l, kfl
    real(c_double) :: v(*), mu(*), mut(*), fd(ip11, 6), tt(*), gx(*), gy(*), gz(*)
    equat = heq
    if (kfl .eq. 2) then
       call met_fludcsa(l, v, mu, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6))
    elseif (kfl .eq. 5) then
       call met_fludko(l, mu, mut, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6))
    else
       call met_fludc(l, v, mu, mut, tt, gx, gy, gz, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6))
    endif
  end subroutine ref_met_flud

  ! sch_turb.f90:4
  subroutine ref_sch_turb(l, u, v, d, fd, ts6, ts7, ncin, sn, lgsnlt, vol, f) bind(C, name="ref_sch_turb")
    use mod_sch_turb
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: u(*), v(*), d(*), fd(ip11, 6), ts6(*), ts7(*), sn(*), vol(*), f(ip00, 6)
    integer(c_int) :: ncin(*)
    call sch_turb(l, u, v, d, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), ts6, ts7, heq, ncin, &
         sn, lgsnlt, vol, f(1,1), f(1,2), f(1,3), f(1,4), f(1,5), f(1,6))
  end subroutine ref_sch_turb

  ! modeleturb%equatt for impli2_eqt_3d's constant cmt (impli2_eqt_3d.f90:82-95)
  subroutine ref_set_equatt(code) bind(C, name="ref_set_equatt")
    use chainecarac
    integer(c_int), value :: code
    select case (code)
    case (0)
       equatt = '1SA   '
    case (1)
       equatt = '2KO   '
    case (2)
       equatt = '2JL   '
    case (3)
       equatt = '2WL   '
    case (4)
       equatt = '2MT   '
    case default
       equatt = '2Sm   '
    end select
  end subroutine ref_set_equatt

  ! impli2_eqt_3d.f90:4 (lparoi = 0)
  subroutine ref_impli2_eqt_3d(l, u, dt, v, mu, mut, cfke, sn, lgsnlt, vol, t1, t2, t3, t4, t5, t6, t7, t8, t9, &
       t10) bind(C, name="ref_impli2_eqt_3d")
    use mod_impli2_eqt_3d
    use chainecarac
    integer(c_int), value :: l, lgsnlt
    real(c_double) :: u(*), dt(*), v(*), mu(*), mut(*), cfke(*), sn(*), vol(*)
    real(c_double) :: t1(*), t2(*), t3(*), t4(*), t5(*), t6(*), t7(*), t8(*), t9(*), t10(*)
    integer :: ncin(1)
    ncin = 0
    equat = heq
    call impli2_eqt_3d(l, u, dt, v, mu, mut, cfke, ncin, 1, sn, lgsnlt, vol, t1, t2, t3, t4, t5, t6, t7, t8, &
         t9, t10)
  end subroutine ref_impli2_eqt_3d

  ! met_samut.f90:4
  subroutine ref_met_samut(l, v, mu, mut) bind(C, name="ref_met_samut")
    use mod_met_samut
    integer(c_int), value :: l
    real(c_double) :: v(*), mu(*), mut(*)
    call met_samut(l, v, mu, mut)
  end subroutine ref_met_samut

  ! met_kocmut.f90:4 (kwsst = 1 computes the velocity gradients itself)
  subroutine ref_met_kocmut(l, iwsst, sn, vol, t, g, dist, v, mu, mut, cmu) bind(C, name="ref_met_kocmut")
    use mod_met_kocmut
    use chainecarac
    integer(c_int), value :: l, iwsst
    real(c_double) :: sn(*), vol(*), t(*), g(ip00, 9), dist(*), v(*), mu(*), mut(*), cmu(*)
    equat = heq
    kwsst = iwsst
    call met_kocmut(l, sn, vol, t, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         dist, v, mu, mut, cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_kocmut

  ! met_cutsa.f90:4 and met_cut.f90:4
  subroutine ref_met_cut(l, which, ek, ee, v) bind(C, name="ref_met_cut")
    use mod_met_cutsa
    use mod_met_cut
    integer(c_int), value :: l, which
    real(c_double), value :: ek, ee
    real(c_double) :: v(*)
    epsk = ek
    epse = ee
    if (which .eq. 2) then
       call met_cutsa(l, v)
    else
       call met_cut(l, v)
    endif
  end subroutine ref_met_cut

  ! met_smkor.f90:4 (sigmd of the model deck)
  subroutine ref_met_smkor(l, sigmdl, v, fd, g, tprod, cfke, ts6, ts7) bind(C, name="ref_met_smkor")
    use mod_met_smkor
    integer(c_int), value :: l
    real(c_double), value :: sigmdl
    real(c_double) :: v(*), fd(ip11, 6), g(ip00, 9), tprod(*), cfke(*), ts6(*), ts7(*)
    sigmd = sigmdl
    call met_smkor(l, v, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), g(1,1), g(1,2), g(1,3), g(1,4), &
         g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), tprod, cfke, ts6, ts7)
  end subroutine ref_met_smkor

  ! met_smch.f90:4
  subroutine ref_met_smch(l, c1, c2, utmin, v, mu, dist, mnpar, utau, tprod, ts6, ts7) bind(C, name="ref_met_smch")
    use mod_met_smch
    integer(c_int), value :: l
    real(c_double), value :: c1, c2, utmin
    real(c_double) :: v(*), mu(*), dist(*), utau(*), tprod(*), ts6(*), ts7(*)
    integer(c_int) :: mnpar(*)
    cke1 = c1
    cke2 = c2
    utaumin = utmin
    call met_smch(l, v, mu, dist, mnpar, utau, tprod, ts6, ts7)
  end subroutine ref_met_smch

  ! Menter constants and switches (atlecdon.f90:255-278)
  subroutine ref_set_menter(mc, kfrac, ksst, imx) bind(C, name="ref_set_menter")
    real(c_double) :: mc(10)
    integer(c_int), value :: kfrac, ksst, imx
    betae = mc(1); okappa = mc(2); sigme1 = mc(3); sigma1 = mc(4); beta1 = mc(5); wsig1 = mc(6)
    sigme2 = mc(7); sigma2 = mc(8); beta2 = mc(9); wsig2 = mc(10)
    kfracom = kfrac
    komsst = ksst
    imxclko = imx
  end subroutine ref_set_menter

  ! Wall faces of module boundary for met_mtcorf1 / atindnor: desc(12, nw)
  ! per wall = block, face (1..6 = i1 i2 j1 j2 k1 k2), imin, imax, jmin,
  ! jmax, kmin, kmax, mmb (points), mpb (offset into ncin), mpn (offset of
  ! the wall's point numbers), unused.  The walls form the list lbdko.
  subroutine ref_set_walls(nw, desc) bind(C, name="ref_set_walls")
    use boundary
    integer(c_int), value :: nw
    integer(c_int) :: desc(12, *)
    character(len=2), parameter :: codes(6) = ['i1', 'i2', 'j1', 'j2', 'k1', 'k2']
    integer :: w
    if (allocated(nfei)) deallocate(nfei, ndlb, indfl, iminb, imaxb, jminb, jmaxb, kminb, kmaxb, mmb, mpb, mpn, lbdko)
    allocate(nfei(nw), ndlb(nw), indfl(nw), iminb(nw), imaxb(nw), jminb(nw), jmaxb(nw), kminb(nw), kmaxb(nw))
    allocate(mmb(nw), mpb(nw), mpn(nw), lbdko(nw))
    do w = 1, nw
       nfei(w) = w
       lbdko(w) = w
       ndlb(w) = desc(1, w)
       indfl(w) = codes(desc(2, w))
       iminb(w) = desc(3, w); imaxb(w) = desc(4, w)
       jminb(w) = desc(5, w); jmaxb(w) = desc(6, w)
       kminb(w) = desc(7, w); kmaxb(w) = desc(8, w)
       mmb(w) = desc(9, w); mpb(w) = desc(10, w); mpn(w) = desc(11, w)
    enddo
    nbdko = nw
  end subroutine ref_set_walls

  ! met_mtcorf1.f90:4
  subroutine ref_met_mtcorf1(l, ncin, dist, mnpar, frac) bind(C, name="ref_met_mtcorf1")
    use mod_met_mtcorf1
    integer(c_int), value :: l
    integer(c_int) :: ncin(*), mnpar(*)
    real(c_double) :: dist(*), frac(*)
    call met_mtcorf1(l, ncin, dist, mnpar, frac)
  end subroutine ref_met_mtcorf1

  ! met_smmt.f90:4
  subroutine ref_met_smmt(l, v, mu, mut, dist, mnpar, ncin, fd, tprod, cfke, gkgo, frac, ts6, ts7) &
       bind(C, name="ref_met_smmt")
    use mod_met_smmt
    integer(c_int), value :: l
    integer(c_int) :: mnpar(*), ncin(*)
    real(c_double) :: v(*), mu(*), mut(*), dist(*), fd(ip11, 6), tprod(*), cfke(*), gkgo(*), frac(*), ts6(*), ts7(*)
    call met_smmt(l, 1, v, mu, mut, dist, mnpar, ncin, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), &
         tprod, cfke, gkgo, frac, ts6, ts7)
  end subroutine ref_met_smmt

  ! met_fludmt.f90:4
  subroutine ref_met_fludmt(l, v, mu, mut, frac, fd) bind(C, name="ref_met_fludmt")
    use mod_met_fludmt
    integer(c_int), value :: l
    real(c_double) :: v(*), mu(*), mut(*), frac(*), fd(ip11, 6)
    call met_fludmt(l, v, mu, mut, frac, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6))
  end subroutine ref_met_fludmt

  ! met_komut.f90:4
  subroutine ref_met_komut(l, sn, vol, t, g, dist, v, mu, mut, cmu) bind(C, name="ref_met_komut")
    use mod_met_komut
    use chainecarac
    integer(c_int), value :: l
    real(c_double) :: sn(*), vol(*), t(*), g(ip00, 9), dist(*), v(*), mu(*), mut(*), cmu(*)
    equat = heq
    call met_komut(l, sn, vol, t, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         dist, v, mu, mut, cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_komut

  ! met_smmtr.f90:4
  subroutine ref_met_smmtr(l, v, mu, mut, dist, mnpar, ncin, fd, g, tprod, cfke, frac, ts6, ts7) &
       bind(C, name="ref_met_smmtr")
    use mod_met_smmtr
    integer(c_int), value :: l
    integer(c_int) :: mnpar(*), ncin(*)
    real(c_double) :: v(*), mu(*), mut(*), dist(*), fd(ip11, 6), g(ip00, 9), tprod(*), cfke(*), frac(*), ts6(*), ts7(*)
    call met_smmtr(l, v, mu, mut, dist, mnpar, ncin, fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), &
         g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), tprod, cfke, frac, ts6, ts7)
  end subroutine ref_met_smmtr

  ! met_smdes.f90:4 (calls met_dist.f90:4, met_difsa.f90:4)
  subroutine ref_met_smdes(l, x, y, z, delta, sn, vol, v, mu, dist, g, fd, cfke, tt, gx, gy, gz, sdif, ts6, ts7, cmu) &
       bind(C, name="ref_met_smdes")
    use mod_met_smdes
    use chainecarac
    integer(c_int), value :: l
    real(c_double) :: x(*), y(*), z(*), delta(*), sn(*), vol(*), v(*), mu(*), dist(*), g(ip00, 9), fd(ip11, 6)
    real(c_double) :: cfke(*), tt(*), gx(*), gy(*), gz(*), sdif(*), ts6(*), ts7(*), cmu(*)
    equat = heq
    call met_smdes(l, x, y, z, delta, sn, vol, v, mu, dist, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), &
         g(1,7), g(1,8), g(1,9), fd(1,1), fd(1,2), fd(1,3), cfke, tt, gx, gy, gz, sdif, ts6, ts7, &
         cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_smdes

  ! met_smsasas.f90:4 (calls met_difsa.f90:4, met_laplaciens.f90:4); isch: the
  ! module's ischema for this call (1: met_laplaciens's plain form)
  subroutine ref_met_smsasas(l, isch, sn, vol, v, mu, dist, g, fd, cfke, tt, gx, gy, gz, sdif, ts6, ts7, tn, cmu) &
       bind(C, name="ref_met_smsasas")
    use mod_met_smsasas
    use chainecarac
    integer(c_int), value :: l, isch
    real(c_double) :: sn(*), vol(*), v(*), mu(*), dist(*), g(ip00, 9), fd(ip11, 6), cfke(*), tt(*), gx(*), gy(*)
    real(c_double) :: gz(*), sdif(*), ts6(*), ts7(*), tn(ip00, 4), cmu(*)
    integer :: keep
    equat = heq
    keep = ischema
    ischema = isch
    call met_smsasas(l, sn, vol, v, mu, dist, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), &
         g(1,9), fd(1,1), fd(1,2), fd(1,3), cfke, tt, gx, gy, gz, sdif, ts6, ts7, tn(1,1), tn(1,2), tn(1,3), &
         tn(1,4), cmu, cmu, cmu, cmu, cmu, cmu)
    ischema = keep
  end subroutine ref_met_smsasas

  ! k-l constants (atlecdon.f90:226-231): cklb1, ckle2, sigmal, xkappa
  subroutine ref_set_kl(kc) bind(C, name="ref_set_kl")
    real(c_double) :: kc(4)
    cklb1 = kc(1); ckle2 = kc(2); sigmal = kc(3); xkappa = kc(4)
  end subroutine ref_set_kl

  ! met_smkl.f90:4 (calls met_bark.f90:4)
  subroutine ref_met_smkl(l, sn, vol, v, mu, mut, dist, mnpar, g, fd, tprod, cfke, tt, gx, gy, gz, bark, ts6, ts7, &
       cmu) bind(C, name="ref_met_smkl")
    use mod_met_smkl
    use chainecarac
    integer(c_int), value :: l
    integer(c_int) :: mnpar(*)
    real(c_double) :: sn(*), vol(*), v(*), mu(*), mut(*), dist(*), g(ip00, 9), fd(ip11, 6), tprod(*), cfke(*)
    real(c_double) :: tt(*), gx(*), gy(*), gz(*), bark(*), ts6(*), ts7(*), cmu(*)
    equat = heq
    call met_smkl(l, sn, vol, v, mu, mut, dist, mnpar, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), &
         g(1,8), g(1,9), fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), tprod, cfke, tt, gx, gy, gz, bark, &
         ts6, ts7, cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_smkl

  ! met_klmut.f90:4
  subroutine ref_met_klmut(l, v, mu, mut, dist) bind(C, name="ref_met_klmut")
    use mod_met_klmut
    integer(c_int), value :: l
    real(c_double) :: v(*), mu(*), mut(*), dist(*)
    call met_klmut(l, v, mu, mut, dist)
  end subroutine ref_met_klmut

  ! k-epsilon constants and switches (atlecdon.f90:112-140): cke1, cke2, cmu;
  ! lparoi; kesst; equatt = '2JL' + the variant letter (code 0 ' ', 1 'S',
  ! 2 'R', 3 'C', 4 'L'), or '2LS' (code 5)
  subroutine ref_set_ke(kc, lpar, kes, code) bind(C, name="ref_set_ke")
    use chainecarac
    real(c_double) :: kc(3)
    integer(c_int), value :: lpar, kes, code
    character(len=1), parameter :: letters(0:4) = (/ ' ', 'S', 'R', 'C', 'L' /)
    cke1 = kc(1); cke2 = kc(2); cmu = kc(3)
    lparoi = lpar
    kesst = kes
    if (code == 5) then
       equatt = '2LS   '
    else
       equatt = '2JL' // letters(code) // '  '
    end if
  end subroutine ref_set_ke

  ! met_smke.f90:4 (met_exgr, met_bark, met_bare, then met_smkes / met_smker
  ! / met_smkec / met_smkesas by equatt(4:4))
  subroutine ref_met_smke(l, isch, sn, vol, v, mu, mut, tprod, cfke, g, tt, gx, gy, gz, bark, bare, ts6, ts7, tn, &
       cson, cmu) bind(C, name="ref_met_smke")
    use mod_met_smke
    use chainecarac
    integer(c_int), value :: l, isch
    real(c_double) :: sn(*), vol(*), v(*), mu(*), mut(*), tprod(*), cfke(*), g(ip00, 9), tt(*), gx(*), gy(*)
    real(c_double) :: gz(*), bark(*), bare(*), ts6(*), ts7(*), tn(ip00, 3), cson(*), cmu(*)
    integer :: keep
    equat = heq
    keep = ischema
    ischema = isch
    call met_smke(l, sn, vol, v, mu, mut, tprod, cfke, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), &
         g(1,8), g(1,9), tt, gx, gy, gz, bark, bare, ts6, ts7, tn(1,1), tn(1,2), tn(1,3), cson, cmu, cmu, cmu, &
         cmu, cmu, cmu)
    ischema = keep
  end subroutine ref_met_smke

  ! met_kemut.f90:4
  subroutine ref_met_kemut(l, sn, vol, v, mu, mut, dist, g, tt, cmu) bind(C, name="ref_met_kemut")
    use mod_met_kemut
    integer(c_int), value :: l
    real(c_double) :: sn(*), vol(*), v(*), mu(*), mut(*), dist(*), g(ip00, 9), tt(*), cmu(*)
    call met_kemut(l, sn, vol, tt, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         dist, v, mu, mut, cmu, cmu, cmu, cmu, cmu, cmu)
  end subroutine ref_met_kemut

  ! met_smklsas.f90:4 (calls met_bark.f90:4 and met_laplaciens.f90:4)
  subroutine ref_met_smklsas(l, isch, sn, vol, v, mu, mut, dist, mnpar, g, fd, tprod, cfke, tt, gx, gy, gz, bark, &
       ts6, ts7, tn, cmu) bind(C, name="ref_met_smklsas")
    use mod_met_smklsas
    use chainecarac
    integer(c_int), value :: l, isch
    integer(c_int) :: mnpar(*)
    real(c_double) :: sn(*), vol(*), v(*), mu(*), mut(*), dist(*), g(ip00, 9), fd(ip11, 6), tprod(*), cfke(*)
    real(c_double) :: tt(*), gx(*), gy(*), gz(*), bark(*), ts6(*), ts7(*), tn(ip00, 4), cmu(*)
    integer :: keep
    equat = heq
    keep = ischema
    ischema = isch
    call met_smklsas(l, sn, vol, v, mu, mut, dist, mnpar, g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), &
         g(1,8), g(1,9), fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), tprod, cfke, tt, gx, gy, gz, bark, &
         ts6, ts7, tn(1,1), tn(1,2), tn(1,3), tn(1,4), cmu, cmu, cmu, cmu, cmu, cmu)
    ischema = keep
  end subroutine ref_met_smklsas

  ! --------------------------------------------------------------------
  ! Boundary tables (module boundary, module mod_mpi on one rank): nfr
  ! frontiers listed one after the other, frontier f holding mt(f) faces
  ! (mpb = mpn = running sum, initbs.f90 / initis.f90); partner(f) is the
  ! coincident frontier of f (0 when f is not coincident); no periodic
  ! rotation (mper = 0).  Frontier numbers are global = local on one rank.
  ! --------------------------------------------------------------------
  subroutine ref_set_boundaries(nfr, mt, partner) bind(C, name="ref_set_boundaries")
    use boundary
    use mod_mpi
    integer(c_int), value :: nfr
    integer(c_int) :: mt(*), partner(*)
    integer :: f, tot
    if (allocated(mmb)) deallocate(mmb, mpb, mpn, lbd, nba, ndcc, mper, ndlb)
    if (allocated(bcl_to_bcg)) deallocate(bcl_to_bcg, bcg_to_proc)
    allocate(mmb(nfr), mpb(nfr), mpn(nfr), lbd(nfr), nba(nfr), ndcc(nfr), mper(nfr), ndlb(nfr))
    allocate(bcl_to_bcg(nfr), bcg_to_proc(nfr))
    tot = 0
    do f = 1, nfr
       mmb(f) = mt(f)
       mpb(f) = tot
       mpn(f) = tot
       tot = tot + mt(f)
       nba(f) = f
       lbd(f) = f
       ndcc(f) = partner(f)
       mper(f) = 0
       ndlb(f) = 1
       bcl_to_bcg(f) = f
       bcg_to_proc(f) = 0
    enddo
    num_bcg = nfr
    nbd = nfr
    ip40 = max(tot, 1)
    ip41 = max(tot, 1)
    ip42 = max(tot, 1)
    ip43 = max(tot, 1)
    ip44 = 1
    protat = 0.
    omg = 0.
  end subroutine ref_set_boundaries

  ! met_bord.f90:4 -- the turbulence ghosts (v(:,6:7)) of every frontier
  ! set by ref_set_boundaries, in frontier order: kinds(f) = 0 coincident
  ! ('rc'), 1 wall ('para'), 2 far field ('nrd'), 3 symmetry ('sym'); kpar:
  ! kparoi; rk, re: rokinf, roeinf.  Coincident frontiers first go through
  ! met_rbvc (the partner frontiers set by ref_set_boundaries).
  subroutine ref_met_bord(nfr, kinds, kpar, rk, re, v, ncbd, ncin) bind(C, name="ref_met_bord")
    use boundary
    use modeleturb
    use mod_met_bord
    integer(c_int), value :: nfr, kpar
    integer(c_int) :: kinds(*), ncbd(*), ncin(*)
    real(c_double), value :: rk, re
    real(c_double) :: v(*)
    character(len=4), parameter :: names(0:3) = ['rc  ', 'para', 'nrd ', 'sym ']
    real(8), allocatable :: bceq(:, :)
    real(8) :: utau(1), mu(1), xnr(1), ynr(1), znr(1)
    integer :: mnc(1), mnr(1), f, nc
    if (allocated(cl)) deallocate(cl)
    if (allocated(nfbc)) deallocate(nfbc)
    allocate(cl(nfr), nfbc(nfr), bceq(ip41, neqt))
    bceq = 0.
    nc = 0
    do f = 1, nfr
       cl(f) = names(kinds(f))
       if (kinds(f) == 0) then
          nc = nc + 1
          nfbc(nc) = f
       endif
    enddo
    mtcx = nc
    mtbx = nfr
    kparoi = kpar
    rokinf = rk
    roeinf = re
    mnc = 0; mnr = 0; utau = 0.; mu = 0.; xnr = 0.; ynr = 0.; znr = 0.
    call met_bord(0, bceq, mnc, ncin, mnr, xnr, ynr, znr, ncbd, v, utau, mu)
    deallocate(bceq)
  end subroutine ref_met_bord

  ! met_parko.f90:4 -- omega wall condition over the walls of ref_set_walls
  ! (equatt '2KO': imxclko = 3, beta1 = 0.075)
  subroutine ref_met_parko(ncin, ncbd, v, mut, mu, dist, mnpar) bind(C, name="ref_met_parko")
    use modeleturb
    use chainecarac
    use mod_met_parko
    integer(c_int) :: ncin(*), ncbd(*), mnpar(*)
    real(c_double) :: v(*), mut(*), mu(*), dist(*)
    equatt = '2KO    '
    call met_parko(0, ncin, ncbd, v, mut, mu, dist, mnpar)
  end subroutine ref_met_parko

  ! the frontier list lbd(1:nl) = fr(1:nl) that rbvc / rbte / rbtc /
  ! met_rbse / met_rbsc walk (zvisqc.f90:182-205, met_num.f90:575-600)
  subroutine set_lbd(nl, fr)
    use boundary
    integer(c_int) :: nl, fr(*)
    integer :: f
    nbd = nl
    do f = 1, nl
       lbd(f) = fr(f)
    enddo
  end subroutine set_lbd

  ! rbvc.f90:4 -- coincident ghosts, face average with the partner frontier
  subroutine ref_rbvc(nl, fr, v, ncbd, ncin, ps, temp, cson) bind(C, name="ref_rbvc")
    use mod_rbvc
    integer(c_int), value :: nl
    integer(c_int) :: fr(*), ncbd(*), ncin(*)
    real(c_double) :: v(*), ps(*), temp(*), cson(*)
    integer :: mnc(1)
    mnc = 0
    call set_lbd(nl, fr)
    call rbvc(v, ncbd, ncin, mnc, ps, temp, cson)
  end subroutine ref_rbvc

  ! residu.f90:4 -- steadiness norms of block l, 5 equations, or 7 with a
  ! transport model (equat(6:7) = 'ke'); kimp = 0 keeps it from printing
  subroutine ref_residu(l, neq, u, v, dt, dumy1, dumy2, dumax, idumx, jdumx, kdumx) bind(C, name="ref_residu")
    use mod_residu
    use chainecarac
    use sortiefichier
    use constantes
    integer(c_int), value :: l, neq
    real(c_double) :: u(*), v(*), dt(*), dumy1(neqt), dumy2(neqt), dumax(neqt)
    integer(c_int) :: idumx(neqt), jdumx(neqt), kdumx(neqt)
    real(c_double), allocatable :: res(:, :)
    integer :: kimp0
    allocate(res(ip00, 7))
    if (neq == 7) then
       equat = 'ns3d ke'
    else
       equat = heq
    endif
    kimp0 = kimp
    kimp = 0
    reelmn = 1.d-300
    call residu(1, l, u, v, dt, res(1,1), res(1,2), res(1,3), res(1,4), res(1,5), res(1,6), res(1,7), 1, &
         dumy1, dumy2, dumax, idumx, jdumx, kdumx)
    kimp = kimp0
    deallocate(res)
  end subroutine ref_residu

  ! rfve.f90:4 -- ghost = adjacent cell (v, p, T, c) on the listed frontiers
  subroutine ref_rfve(nl, fr, v, ps, temp, cson, ncbd, ncin) bind(C, name="ref_rfve")
    use mod_rfve
    integer(c_int), value :: nl
    integer(c_int) :: fr(*), ncbd(*), ncin(*)
    real(c_double) :: v(*), ps(*), temp(*), cson(*)
    call set_lbd(nl, fr)
    call rfve(v, ps, temp, cson, ncbd, ncin)
  end subroutine ref_rfve

  ! clpara.f90:4 -- adiabatic wall, one frontier
  subroutine ref_clpara(mfb, ncbd, v, ncin, ps, temp, cson) bind(C, name="ref_clpara")
    use boundary
    use mod_clpara
    integer(c_int), value :: mfb
    integer(c_int) :: ncbd(*), ncin(*)
    real(c_double) :: v(*), ps(*), temp(*), cson(*)
    call clpara(mfb, ncbd, v, mmb, mpb, ncin, ps, temp, cson)
  end subroutine ref_clpara

  ! rbord.f90:300-323 -- 'nrd' frontier of block l: utnrd.f90:4 (state from
  ! Mach, incidence angles alpha, beta and the stagnation state roa1, aa1,
  ! pa1), rbve.f90:4 (ghost = adjacent cell), clnrd.f90:4.  w0(5) returns
  ! the state of the frontier's first face.
  subroutine ref_farfield(mfb, l, mach, alpha, beta, roa, aa, pa, nxn, nyn, nzn, ncbd, ncin, v, ps, temp, cson, &
       y, z, w0) bind(C, name="ref_farfield")
    use boundary
    use definition
    use mod_utnrd
    use mod_rbve
    use mod_clnrd
    integer(c_int), value :: mfb, l
    real(c_double), value :: mach, alpha, beta, roa, aa, pa
    integer(c_int) :: ncbd(*), ncin(*)
    real(c_double) :: nxn(*), nyn(*), nzn(*), v(*), ps(*), temp(*), cson(*), y(*), z(*), w0(5)
    real(c_double), allocatable :: bceqt(:, :), rod(:), roud(:), rovd(:), rowd(:), roed(:)
    integer :: nl, fr(1)
    allocate(bceqt(ip41, neqt), rod(ip40), roud(ip40), rovd(ip40), rowd(ip40), roed(ip40))
    bceqt = 0.
    bceqt(mpb(mfb) + 1, 1) = mach
    bceqt(mpb(mfb) + 1, 2) = alpha
    bceqt(mpb(mfb) + 1, 3) = beta
    roa1 = roa
    aa1 = aa
    pa1 = pa
    call utnrd(bceqt, mfb, l, rod, roud, rovd, rowd, roed, ncbd, y, z)
    nl = 1
    fr(1) = mfb
    call set_lbd(nl, fr)
    call rbve(v, ncin, ncbd)
    call clnrd(mfb, rod, roud, rovd, rowd, roed, nxn, nyn, nzn, ncbd, v, mmb, mpb, mpn, l, ps, temp, cson)
    w0(1) = rod(1)
    w0(2) = roud(1)
    w0(3) = rovd(1)
    w0(4) = rowd(1)
    w0(5) = roed(1)
    deallocate(bceqt, rod, roud, rovd, rowd, roed)
  end subroutine ref_farfield

  ! clsym.f90:4 -- symmetry plane, one frontier
  subroutine ref_clsym(mfb, nxn, nyn, nzn, ncin, ncbd, v) bind(C, name="ref_clsym")
    use boundary
    use mod_clsym
    integer(c_int), value :: mfb
    integer(c_int) :: ncin(*), ncbd(*)
    real(c_double) :: nxn(*), nyn(*), nzn(*), v(*)
    call clsym(mfb, nxn, nyn, nzn, ncin, ncbd, v, mmb, mpb, mpn)
  end subroutine ref_clsym

  ! rbte.f90:4 -- stress and heat-flux ghosts = adjacent cell (tq: 9 planes)
  subroutine ref_rbte(nl, fr, tq, ncbd, ncin) bind(C, name="ref_rbte")
    use mod_rbte
    integer(c_int), value :: nl
    integer(c_int) :: fr(*), ncbd(*), ncin(*)
    real(c_double) :: tq(ip12, 9)
    call set_lbd(nl, fr)
    call rbte(tq(1,1), tq(1,2), tq(1,3), tq(1,4), tq(1,5), tq(1,6), tq(1,7), tq(1,8), tq(1,9), ncbd, ncin)
  end subroutine ref_rbte

  ! rbtc.f90:4 -- coincident stress and heat-flux ghosts
  subroutine ref_rbtc(nl, fr, tq, ncbd, ncin) bind(C, name="ref_rbtc")
    use mod_rbtc
    integer(c_int), value :: nl
    integer(c_int) :: fr(*), ncbd(*), ncin(*)
    real(c_double) :: tq(ip12, 9)
    integer :: mnc(1)
    mnc = 0
    call set_lbd(nl, fr)
    call rbtc(tq(1,1), tq(1,2), tq(1,3), tq(1,4), tq(1,5), tq(1,6), tq(1,7), tq(1,8), tq(1,9), ncbd, ncin, mnc)
  end subroutine ref_rbtc

  ! zvisqc.f90:4 -- stresses and heat flux of every block (zgrad2, then
  ! zfluto), then their ghosts: rbte over the non-coincident frontiers
  ! frn(1:nn) (the list nfbn), rbtc over the coincident ones frc(1:nc) (the
  ! list nfbc); no adjacent-frontier list (mtax = 0).  icyc = 0 keeps the
  ! transition and multigrid branches off; the gradient weights cmui1..cmuk2
  ! are 1 (inivec.f90:107-115); ischema = 7 selects zgrad2; equat (module
  ! chainecarac, which zvisqc reads) is the 3-D Navier-Stokes set.
  subroutine ref_zvisqc(nn, frn, nc, frc, s, mu, mut, tq, ncbd, ncin, sn, vol, temp) bind(C, name="ref_zvisqc")
    use boundary
    use chainecarac
    use mod_zvisqc
    integer(c_int), value :: nn, nc
    integer(c_int) :: frn(*), frc(*), ncbd(*), ncin(*)
    real(c_double) :: s(*), mu(*), mut(*), tq(ip12, 9), sn(*), vol(*), temp(*)
    ! the routine's scratch and unused inputs, kept between calls (as a
    ! driver keeps its arrays) and remade when the layout grows
    real(c_double), allocatable, save :: xyz(:), cmu(:), g(:, :), dist(:), fgam(:)
    integer, allocatable, save :: mnpar(:)
    integer :: mnc(1), f, isch0
    if (allocated(xyz)) then
       if (size(xyz) < ip21 .or. size(g, 1) < ip00 .or. size(dist) < ip12) &
            deallocate(xyz, cmu, g, dist, fgam, mnpar)
    endif
    if (.not. allocated(xyz)) then
       allocate(xyz(ip21), cmu(ip21), g(ip00, 9), dist(ip12), fgam(ip12), mnpar(ip12))
       xyz = 0.; cmu = 1.; g = 0.; dist = 1.; fgam = 0.; mnpar = 0
    endif
    mnc = 0
    if (allocated(nfbn)) deallocate(nfbn)
    if (allocated(nfbc)) deallocate(nfbc)
    if (allocated(nfba)) deallocate(nfba)
    allocate(nfbn(max(nn, 1)), nfbc(max(nc, 1)), nfba(1))
    do f = 1, nn
       nfbn(f) = frn(f)
    enddo
    do f = 1, nc
       nfbc(f) = frc(f)
    enddo
    mtnx = nn; mtcx = nc; mtax = 0; mtb = 0
    isch0 = ischema
    ischema = 7
    equat = heq
    call zvisqc(1, s, mu, s, xyz, xyz, xyz, mut, dist, mnpar, fgam, &
         tq(1,1), tq(1,2), tq(1,3), tq(1,4), tq(1,5), tq(1,6), tq(1,7), tq(1,8), tq(1,9), &
         0, 0, ncbd, ncin, mnc, sn, vol, &
         g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,6), g(1,7), g(1,8), g(1,9), &
         cmu, cmu, cmu, cmu, cmu, cmu, temp)
    ischema = isch0
  end subroutine ref_zvisqc

  ! met_rbse.f90:4 -- turbulence flux-density ghosts = adjacent cell (fd: 6 planes)
  subroutine ref_met_rbse(nl, fr, fd, ncbd, ncin) bind(C, name="ref_met_rbse")
    use mod_met_rbse
    integer(c_int), value :: nl
    integer(c_int) :: fr(*), ncbd(*), ncin(*)
    real(c_double) :: fd(ip12, 6)
    call set_lbd(nl, fr)
    call met_rbse(fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), ncbd, ncin)
  end subroutine ref_met_rbse

  ! met_rbsc.f90:4 -- coincident turbulence flux-density ghosts
  subroutine ref_met_rbsc(nl, fr, fd, ncbd, ncin) bind(C, name="ref_met_rbsc")
    use mod_met_rbsc
    integer(c_int), value :: nl
    integer(c_int) :: fr(*), ncbd(*), ncin(*)
    real(c_double) :: fd(ip12, 6)
    integer :: mnc(1)
    mnc = 0
    call set_lbd(nl, fr)
    call met_rbsc(fd(1,1), fd(1,2), fd(1,3), fd(1,4), fd(1,5), fd(1,6), ncbd, ncin, mnc)
  end subroutine ref_met_rbsc

  ! initns.f90:4 -- inward unit normals of frontier mfb, face `iface`
  ! (1..6: i1 i2 j1 j2 k1 k2) of block l over the node range imin..kmax
  subroutine ref_initns(mfb, l, iface, imin, imax, jmin, jmax, kmin, kmax, x, y, z, ncbd, nxn, nyn, nzn) &
       bind(C, name="ref_initns")
    use mod_initns
    integer(c_int), value :: mfb, l, iface, imin, imax, jmin, jmax, kmin, kmax
    integer(c_int) :: ncbd(*)
    real(c_double) :: x(*), y(*), z(*), nxn(*), nyn(*), nzn(*)
    real(c_double), allocatable :: tn(:, :)
    character(len=2), parameter :: fname(6) = (/ 'i1', 'i2', 'j1', 'j2', 'k1', 'k2' /)
    allocate(tn(ip00, 9))
    tn = 0.
    call initns(mfb, l, fname(iface), imin, imax, jmin, jmax, kmin, kmax, heq, x, y, z, ncbd, nxn, nyn, nzn, &
         tn(1,1), tn(1,2), tn(1,3), tn(1,4), tn(1,5), tn(1,6), tn(1,7), tn(1,8), tn(1,9))
    deallocate(tn)
  end subroutine ref_initns

  ! --------------------------------------------------------------------
  ! On-disk formats: the reference's writda / readda (restart) and writdg /
  ! readdg (mesh), every block in order on one sequential unformatted unit,
  ! as svfw.f90:58 and infw.f90:69 call them block after block.  The
  ! ordering helpers of mod_mpi take block l as its own global block on
  ! rank 0 (the serial build).
  ! --------------------------------------------------------------------
  subroutine serial_blocks(nb)
    use mod_mpi
    integer, intent(in) :: nb
    integer :: l
    if (allocated(bl_to_bg)) deallocate(bl_to_bg)
    if (allocated(bg_to_proc)) deallocate(bg_to_proc)
    allocate(bl_to_bg(nb), bg_to_proc(nb))
    do l = 1, nb
       bl_to_bg(l) = l
       bg_to_proc(l) = 0
    enddo
  end subroutine serial_blocks

  function cstr(p, n) result(fn)
    integer(c_int), intent(in) :: n
    character(kind=c_char), intent(in) :: p(n)
    character(len=n) :: fn
    integer :: i
    do i = 1, n
       fn(i:i) = p(i)
    enddo
  end function cstr

  ! writda.f90:4 with chdacc 'cccc' extents (the real cells); eq: equat
  subroutine ref_writda_all(path, plen, nb, eq, v, mut) bind(C, name="ref_writda_all")
    use chainecarac
    use mod_writda
    integer(c_int), value :: plen, nb
    character(kind=c_char) :: path(plen), eq(7)
    real(c_double) :: v(ip11, *), mut(*)
    real(8) :: utau(1)
    integer :: l, n0
    call serial_blocks(nb)
    utau = 0.
    open(unit=31, file=cstr(path, plen), form='unformatted', access='sequential', status='replace')
    do l = 1, nb
       n0 = npc(l)
       call writda(l, 31, cstr(eq, 7), utau, ii1(l), ii2(l) - 1, jj1(l), jj2(l) - 1, kk1(l), kk2(l) - 1, &
            v(n0 + 1, 1), v(n0 + 1, 2), v(n0 + 1, 3), v(n0 + 1, 4), v(n0 + 1, 5), v(n0 + 1, 6), v(n0 + 1, 7), &
            ip12, mut(n0 + 1))
    enddo
    close(31)
  end subroutine ref_writda_all

  ! readda.f90:4 (kfmg = 0, kutau = 0: no dual-time copies, no utau)
  subroutine ref_readda_all(path, plen, nb, eq, v, mut) bind(C, name="ref_readda_all")
    use chainecarac
    use mod_readda
    integer(c_int), value :: plen, nb
    character(kind=c_char) :: path(plen), eq(7)
    real(c_double) :: v(ip11, ip60), mut(ip12)
    real(8), allocatable :: utau(:)
    integer :: l
    call serial_blocks(nb)
    equat = cstr(eq, 7)
    kfmg = 0
    kutau = 0
    keinit = 0
    allocate(utau(1))
    open(unit=32, file=cstr(path, plen), form='unformatted', access='sequential', status='old')
    do l = 1, nb
       call readda(l, 32, v, mut, utau, v, v, v)
    enddo
    close(32)
  end subroutine ref_readda_all

  ! writdg.f90:4 over the nodes ii1..ii2, jj1..jj2, kk1..kk2
  subroutine ref_writdg_all(path, plen, nb, x, y, z) bind(C, name="ref_writdg_all")
    use mod_writdg
    integer(c_int), value :: plen, nb
    character(kind=c_char) :: path(plen)
    real(c_double) :: x(*), y(*), z(*)
    integer :: l, n0
    call serial_blocks(nb)
    open(unit=33, file=cstr(path, plen), form='unformatted', access='sequential', status='replace')
    do l = 1, nb
       n0 = npn(l)
       call writdg(l, 33, ii1(l), ii2(l), jj1(l), jj2(l), kk1(l), kk2(l), x(n0 + 1), y(n0 + 1), z(n0 + 1))
    enddo
    close(33)
  end subroutine ref_writdg_all

  ! readdg.f90:4
  subroutine ref_readdg_all(path, plen, nb, x, y, z) bind(C, name="ref_readdg_all")
    use mod_readdg
    integer(c_int), value :: plen, nb
    character(kind=c_char) :: path(plen)
    real(c_double) :: x(ip21), y(ip21), z(ip21)
    integer :: l
    call serial_blocks(nb)
    open(unit=34, file=cstr(path, plen), form='unformatted', access='sequential', status='old')
    do l = 1, nb
       call readdg(l, 34, x, y, z)
    enddo
    close(34)
  end subroutine ref_readdg_all

end module codens_ref_harness
