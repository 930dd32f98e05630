! Drop-in replacements, on the GPU, for the per-block routines of the
! residual-and-update path.  Each module keeps the reference's module name,
! subroutine name and dummy-argument list, so the reference's callers
! (atsch_num.f90, met_num.f90, smg_flu.f90) compile against them unchanged
! when these files are built in place of the originals.
!
! Every call moves block lm's slice of the arrays it reads to the device
! (codens_dev), runs the C-ABI entry point of include/codens.h on the
! library's stream, and copies back the slice of the arrays it writes
! (codens_back).  Output arrays are copied in as well, so entries a kernel
! leaves alone keep the caller's values.  Host scratch arguments (rhol..prr,
! tn*, dfxx..rv, sfsi.., snu, del6/7) are not touched: the kernels keep their
! intermediates on the device.  Only the 3-D form is provided
! (equat(3:4) = '3d'); other forms stop.
!
! Set codens_geom_once = .true. (module codens_c) once all metrics sn and
! volumes vol are final to upload them only on first use.  With
! codens_host_residency(1) the copies back are deferred and the copies in
! skipped while the device holds the current values: the driver then calls
! codens_sync before its own code reads an array and codens_modified after
! its own code writes one (include/codens.h).

module mod_zpres
  implicit none
contains
  ! zpres.f90:4 -- pressure, temperature and sound speed of block l.
  subroutine zpres(l, v, pression, ztemp, cson)
    use para_var
    use codens_c
    implicit none
    integer :: l
    double precision, target :: v(ip11,ip60), pression(ip11), ztemp(ip11), cson(ip11)
    type(codens_block) :: b
    type(c_ptr) :: dv, dp, dtm, dc
    b = codens_block_of(l)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
    dp = codens_dev(c_loc(pression), ip11, ip11, 1, 1, b, 1)
    dtm = codens_dev(c_loc(ztemp), ip11, ip11, 1, 1, b, 1)
    dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_zpres(b, codens_phys_now(), dv, int(ip11, c_int64_t), dp, dtm, dc, c_null_ptr), &
         'zpres')
    call codens_back(c_loc(pression), ip11, 1, 1, b)
    call codens_back(c_loc(ztemp), ip11, 1, 1, b)
    call codens_back(c_loc(cson), ip11, 1, 1, b)
  end subroutine zpres
end module mod_zpres

module mod_zvismo
  implicit none
contains
  ! zvismo.f90:4 -- molecular viscosity; stops on T <= 0 as the reference
  ! does (zvismo.f90:122-127).
  subroutine zvismo(l, mu, s, temp)
    use para_var
    use sortiefichier
    use codens_c
    implicit none
    integer :: l
    double precision, target :: mu(ip12), s(ip11,ip60), temp(ip11)
    integer(c_int32_t), target, save :: nbad
    type(codens_block) :: b
    type(c_ptr) :: dmu, dtm, dn
    b = codens_block_of(l)
    dtm = codens_dev(c_loc(temp), ip11, ip11, 1, 1, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    nbad = 0
    call codens_modified(c_loc(nbad))
    call codens_check(codens_host_mirror(c_loc(nbad), 4_c_int64_t, 0_c_int64_t, 4_c_int64_t, 4_c_int64_t, &
         1_c_int32_t, 1_c_int32_t, dn, c_null_ptr), 'zvismo')
    call codens_check(codens_zvismo(b, codens_phys_now(), dtm, dmu, dn, c_null_ptr), 'zvismo')
    call codens_back(c_loc(mu), ip12, 1, 1, b)
    call codens_check(codens_host_fetch(c_loc(nbad), 0_c_int64_t, 4_c_int64_t, 4_c_int64_t, 1_c_int32_t, &
         c_null_ptr), 'zvismo')
    call codens_sync(c_loc(nbad))
    if (nbad .ne. 0) then
       write(imp,'(/,"!!!zvismo: temperature negative en",i8," cellules domaine= ",i4)') nbad, l
       error stop 'zvismo: non-positive temperature'
    endif
  end subroutine zvismo
end module mod_zvismo

module mod_zvisqc
  use iso_c_binding, only: c_int64_t, c_ptr, c_null_ptr
  implicit none
  ! device copies of the stress-ghost lists (0-based storage indices): rbte's
  ! (ghost, adjacent cell) over nfbn and nfba, rbtc's (ghost, adjacent cell,
  ! cell across) over nfbc, kept while they stay the same
  integer(c_int64_t), allocatable, target, save :: zq_g(:), zq_i(:), zq_cg(:), zq_ci(:), zq_cr(:)
  ! the gradient-weight array last found all ones (checked once per array)
  type(c_ptr), save :: zq_cmu = c_null_ptr
contains
  ! zvisqc.f90:4 -- stresses and heat flux toxx..qcz of every block of grid
  ! level img: Green-Gauss gradients (zgrad2) and the stress and heat-flux
  ! closure (zfluto) in one device pass per block, then the ghosts of
  ! zvisqc.f90:182-205 on the device: rbte's copies over the frontier lists
  ! nfbn and nfba, rbtc's averages over nfbc.  Where rbtc has no device form
  ! (a rotated periodic frontier, a partner frontier on another process or
  ! outside the list) the reference's own rbte and rbtc run on the host.  The
  ! transition and multigrid updates of mut (atctranske, smg_fcs) run on the
  ! host first, as they read neither the gradients nor the stresses.  The
  ! gradient scratch dvxx..dvzz is not written.  Under residency, host
  ! routines' arrays are synchronised before them and marked written after
  ! them.  Stops for ischema = 1 (zgrad's gradients) and for gradient weights
  ! cmui1..cmuk2 other than 1 (inivec.f90:107-115; checked on the first call
  ! with a given cmui1, as the reference sets them once).
  subroutine zvisqc(img, s, mu, ro, x, y, z, mut, dist, mnpar, fgam, toxx, toxy, toxz, toyy, toyz, tozz, &
       qcx, qcy, qcz, icyc, mcyturb, ncbd, ncin, mnc, sn, vol, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, &
       dvzy, dvzz, cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2, ztemp)
    use para_var
    use para_fige
    use boundary
    use maillage
    use chainecarac
    use schemanum
    use modeleturb
    use mod_smg_fcs
    use mod_atctranske
    use mod_rbte
    use mod_rbtc
    use codens_c
    implicit none
    integer :: icyc, img, mcyturb, mnc(ip43), mnpar(ip12), ncbd(ip41), ncin(ip41)
    double precision, target :: s(ip11,ip60), mu(ip12), mut(ip12), sn(ip31*ndir), vol(ip11), ztemp(ip11)
    double precision, target :: toxx(ip12), toxy(ip12), toxz(ip12), toyy(ip12), toyz(ip12), tozz(ip12)
    double precision, target :: qcx(ip12), qcy(ip12), qcz(ip12)
    double precision :: ro(ip11), x(ip21), y(ip21), z(ip21), dist(ip12), fgam(ip12)
    double precision :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision, target :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dv, dtm, dmu, dmut, dsn, dvol, dtq, tq(9)
    integer :: k, l, lm, mf, npsn, n0, n1
    logical :: check_cmu
    call codens_need_3d(equat, 'zvisqc')
    if (ischema .eq. 1) error stop 'zvisqc: ischema = 1 (zgrad) has no device form'
    tq = [c_loc(toxx), c_loc(toxy), c_loc(toxz), c_loc(toyy), c_loc(toyz), c_loc(tozz), c_loc(qcx), &
         c_loc(qcy), c_loc(qcz)]
    dtq = c_null_ptr
    check_cmu = .not. c_associated(zq_cmu, c_loc(cmui1))
    do l = 1, lzx
       lm = l + (img-1)*lz
       n0 = npc(lm) + 1
       n1 = npc(lm) + nnn(lm)
       if (check_cmu) then
          if (any(cmui1(n0:n1) .ne. 1.) .or. any(cmui2(n0:n1) .ne. 1.) .or. any(cmuj1(n0:n1) .ne. 1.) .or. &
              any(cmuj2(n0:n1) .ne. 1.) .or. any(cmuk1(n0:n1) .ne. 1.) .or. any(cmuk2(n0:n1) .ne. 1.)) &
              error stop 'zvisqc: gradient weights other than 1 have no device form'
       endif
       if (icyc .gt. icytur0 .and. mcyturb .eq. 1) then
          call codens_sync(c_loc(s)); call codens_sync(c_loc(mu)); call codens_sync(c_loc(mut))
          call codens_sync(c_loc(vol))
          if (img .eq. mgl) then
             if (ktransi .eq. 1) call atctranske(l, s, mu, mut, mnpar, fgam)
          else
             call smg_fcs(img-1, img, vol, mut)
          endif
          call codens_modified(c_loc(mut))
       endif
       b = codens_block_of(lm)
       npsn = ndir*npfb(lm) + 1
       dv = codens_dev(c_loc(s), ip11*ip60, ip11, 1, 5, b, 1)
       dtm = codens_dev(c_loc(ztemp), ip11, ip11, 1, 1, b, 1)
       dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
       dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
       dsn = codens_dev_all(c_loc(sn(npsn)), 9*nnn(lm), codens_geom_mode())
       dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
       dtq = codens_tq_in(tq, ip12, b)
       call codens_check(codens_zvisqc(b, codens_phys_now(), dsn, dvol, dv, int(ip11, c_int64_t), dtm, dmu, &
            dmut, dtq, c_null_ptr), 'zvisqc')
    enddo
    zq_cmu = c_loc(cmui1)
    ! the ghosts: on the device when every listed frontier has a device form
    ! (the bundle of all blocks' stresses is dtq), else on the host
    if (c_associated(dtq) .and. ghost_lists(img, ncbd, ncin)) then
       if (size(zq_g) > 0) call codens_check(codens_bc_tq_copy(int(size(zq_g), c_int64_t), &
            list_dev(zq_g), list_dev(zq_i), dtq, int(ip12, c_int64_t), c_null_ptr), 'zvisqc')
       if (size(zq_cg) > 0) call codens_check(codens_bc_tq_coincident(int(size(zq_cg), c_int64_t), &
            list_dev(zq_cg), list_dev(zq_ci), list_dev(zq_cr), dtq, int(ip12, c_int64_t), c_null_ptr), 'zvisqc')
       do l = 1, lzx
          call codens_tq_back(tq, ip12, codens_block_of(l + (img-1)*lz))
       enddo
       return
    endif
    do l = 1, lzx
       call codens_tq_back(tq, ip12, codens_block_of(l + (img-1)*lz))
    enddo
    do k = 1, 9
       call codens_sync(tq(k))
    enddo
    do mf = 1, mtnx
       lbd(mf) = nfbn(mf) + (img-1)*mtb
    enddo
    nbd = mtnx
    call rbte(toxx, toxy, toxz, toyy, toyz, tozz, qcx, qcy, qcz, ncbd, ncin)
    do mf = 1, mtax
       lbd(mf) = nfba(mf) + (img-1)*mtb
    enddo
    nbd = mtax
    call rbte(toxx, toxy, toxz, toyy, toyz, tozz, qcx, qcy, qcz, ncbd, ncin)
    do mf = 1, mtcx
       lbd(mf) = nfbc(mf) + (img-1)*mtb
    enddo
    nbd = mtcx
    call rbtc(toxx, toxy, toxz, toyy, toyz, tozz, qcx, qcy, qcz, ncbd, ncin, mnc)
    do k = 1, 9
       call codens_modified(tq(k))
    enddo
  end subroutine zvisqc

  ! device address of a ghost list (copied when first made or rebuilt)
  function list_dev(a) result(dev)
    use codens_c
    integer(c_int64_t), target, intent(in) :: a(:)
    type(c_ptr) :: dev
    integer(c_int64_t) :: nb
    nb = 8_c_int64_t * size(a)
    call codens_check(codens_host_mirror(c_loc(a), nb, 0_c_int64_t, nb, nb, 1_c_int32_t, 1_c_int32_t, dev, &
         c_null_ptr), 'zvisqc')
  end function list_dev

  ! The stress-ghost lists of rbte over nfbn, nfba and of rbtc over nfbc
  ! (zvisqc.f90:182-205), face m of frontier f at mpb(f)+m: ghost ncbd,
  ! adjacent cell ncin, and for rbtc the partner frontier's adjacent cell of
  ! the same face number (rbtc.f90:54-117).  Built on every call and kept
  ! (with their device copies) while they come out the same.  False where
  ! rbtc has no device form: a rotated periodic frontier (mper /= 0), a
  ! partner on another process, outside nfbc, or of another face count.
  logical function ghost_lists(img, ncbd, ncin)
    use para_var
    use para_fige
    use boundary
    use mod_mpi
    use codens_c
    integer, intent(in) :: img, ncbd(ip41), ncin(ip41)
    integer(c_int64_t), allocatable :: g(:), i(:), cg(:), ci(:), cr(:)
    integer, allocatable :: fr(:), part(:)
    integer :: mf2, mfb, m, nn, nc, q, f, nlist
    ghost_lists = .false.
    nlist = mtnx + mtax + mtcx
    allocate(fr(nlist), part(nlist))
    fr(1:mtnx) = nfbn(1:mtnx) + (img-1)*mtb
    fr(mtnx+1:mtnx+mtax) = nfba(1:mtax) + (img-1)*mtb
    fr(mtnx+mtax+1:) = nfbc(1:mtcx) + (img-1)*mtb
    part = 0
    nn = 0
    do q = 1, mtnx + mtax
       nn = nn + mmb(fr(q))
    enddo
    nc = 0
    do q = mtnx + mtax + 1, nlist
       mfb = fr(q)
       if (mper(mfb) /= 0) return
       do mf2 = mtnx + mtax + 1, nlist
          if (bcl_to_bcg(fr(mf2)) == ndcc(mfb)) part(q) = fr(mf2)
       enddo
       if (part(q) == 0) return
       if (bcg_to_proc(bcl_to_bcg(mfb)) /= bcg_to_proc(ndcc(mfb)) .or. mmb(part(q)) /= mmb(mfb)) return
       nc = nc + mmb(mfb)
    enddo
    allocate(g(nn), i(nn), cg(nc), ci(nc), cr(nc))
    f = 0
    do q = 1, mtnx + mtax
       do m = 1, mmb(fr(q))
          f = f + 1
          g(f) = ncbd(mpb(fr(q)) + m) - 1
          i(f) = ncin(mpb(fr(q)) + m) - 1
       enddo
    enddo
    f = 0
    do q = mtnx + mtax + 1, nlist
       do m = 1, mmb(fr(q))
          f = f + 1
          cg(f) = ncbd(mpb(fr(q)) + m) - 1
          ci(f) = ncin(mpb(fr(q)) + m) - 1
          cr(f) = ncin(mpb(part(q)) + m) - 1
       enddo
    enddo
    ghost_lists = .true.
    if (allocated(zq_g)) then
       if (same(zq_g, g) .and. same(zq_i, i) .and. same(zq_cg, cg) .and. same(zq_ci, ci) .and. same(zq_cr, cr)) &
            return
       ! other lists: the device copies of the kept ones are stale
       call forget(zq_g); call forget(zq_i); call forget(zq_cg); call forget(zq_ci); call forget(zq_cr)
       deallocate(zq_g, zq_i, zq_cg, zq_ci, zq_cr)
    endif
    call move_alloc(g, zq_g); call move_alloc(i, zq_i)
    call move_alloc(cg, zq_cg); call move_alloc(ci, zq_ci); call move_alloc(cr, zq_cr)
  contains
    subroutine forget(a)
      integer(c_int64_t), target, intent(in) :: a(:)
      if (size(a) > 0) call codens_modified(c_loc(a))
    end subroutine forget

    logical function same(a, b)
      integer(c_int64_t), intent(in) :: a(:), b(:)
      same = size(a) == size(b)
      if (same) same = all(a == b)
    end function same
  end function ghost_lists
end module mod_zvisqc

module mod_chronos
  implicit none
contains
  ! chronos.f90:4 -- local time step of block l (viscous limit for 'ns').
  subroutine chronos(l, etal, mu, mut, t, dt, equat, sn, lgsnlt, vol, sfsi, sfsj, sfsk, dism, cson)
    use para_var
    use para_fige
    use codens_c
    implicit none
    integer :: l, lgsnlt
    double precision :: etal
    double precision, target :: mu(ip12), mut(ip12), t(ip11,ip60), dt(ip11), sn(lgsnlt,nind,ndir), vol(ip11)
    double precision, target :: cson(ip11)
    double precision :: sfsi(ip00), sfsj(ip00), sfsk(ip00), dism(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: dmu, dmut, dv, ddt, dsn, dvol, dc
    integer(c_int32_t) :: ns
    call codens_need_3d(equat, 'chronos')
    b = codens_block_of(l)
    ns = 0
    if (equat(1:2) .eq. 'ns') ns = 1
    dv = codens_dev(c_loc(t), ip11*ip60, ip11, 1, 5, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddt = codens_dev(c_loc(dt), ip11, ip11, 1, 1, b, 1)
    dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    call codens_check(codens_chronos(b, codens_phys_now(), etal, ns, dv, int(ip11, c_int64_t), dmu, dmut, &
         dsn, dvol, dc, ddt, c_null_ptr), 'chronos')
    call codens_back(c_loc(dt), ip11, 1, 1, b)
  end subroutine chronos
end module mod_chronos

module mod_sch_roe_euler
  implicit none
contains
  ! sch_roe_euler.f90:4 -- Euler Roe residual u of block lm (and ff -= u
  ! when ityprk /= 0).
  subroutine sch_roe_euler(lm, ityprk, u, v, ff, equat, sn, lgsnlt, rhol, ul, vl, wl, pl, rhor, ur, vr, wr, &
       prr, ps)
    use para_var
    use para_fige
    use codens_c
    implicit none
    integer :: lm, ityprk, lgsnlt
    double precision, target :: u(ip11,ip60), v(ip11,ip60), ff(ip11,ip60), sn(lgsnlt,nind,ndir), ps(ip11)
    double precision :: rhol(ip00), ul(ip00), vl(ip00), wl(ip00), pl(ip00)
    double precision :: rhor(ip00), ur(ip00), vr(ip00), wr(ip00), prr(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: du, dv, dff, dsn, dps
    call codens_need_3d(equat, 'sch_roe_euler')
    b = codens_block_of(lm)
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 1, 5, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
    dff = c_null_ptr
    if (ityprk .ne. 0) dff = codens_dev(c_loc(ff), ip11*ip60, ip11, 1, 5, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dps = codens_dev(c_loc(ps), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_sch_roe_euler(b, codens_phys_now(), codens_scheme_now(), int(ityprk, c_int32_t), &
         du, dv, int(ip11, c_int64_t), dff, dsn, dps, c_null_ptr), 'sch_roe_euler')
    call codens_back(c_loc(u), ip11, 1, 5, b)
    if (ityprk .ne. 0) call codens_back(c_loc(ff), ip11, 1, 5, b)
  end subroutine sch_roe_euler
end module mod_sch_roe_euler

module mod_sch_roe
  implicit none
contains
  ! sch_roe.f90:4 -- Navier-Stokes Roe residual u of block lm: MUSCL with the
  ! limiter of module schemanum, plus the viscous face terms of toxx..qcz.
  subroutine sch_roe(lm, ityprk, u, v, ff, toxx, toxy, toxz, toyy, toyz, tozz, qcx, qcy, qcz, equat, sn, &
       lgsnlt, rhol, ul, vl, wl, pl, rhor, ur, vr, wr, prr, ps)
    use para_var
    use para_fige
    use codens_c
    implicit none
    integer :: lm, ityprk, lgsnlt
    double precision, target :: u(ip11,ip60), v(ip11,ip60), ff(ip11,ip60), sn(lgsnlt,nind,ndir), ps(ip11)
    double precision, target :: toxx(ip12), toxy(ip12), toxz(ip12), toyy(ip12), toyz(ip12), tozz(ip12)
    double precision, target :: qcx(ip12), qcy(ip12), qcz(ip12)
    double precision :: rhol(ip00), ul(ip00), vl(ip00), wl(ip00), pl(ip00)
    double precision :: rhor(ip00), ur(ip00), vr(ip00), wr(ip00), prr(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: du, dv, dff, dtq, dsn, dps, tq(9)
    call codens_need_3d(equat, 'sch_roe')
    b = codens_block_of(lm)
    tq = [c_loc(toxx), c_loc(toxy), c_loc(toxz), c_loc(toyy), c_loc(toyz), c_loc(tozz), c_loc(qcx), &
         c_loc(qcy), c_loc(qcz)]
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 1, 5, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
    dff = c_null_ptr
    if (ityprk .ne. 0) dff = codens_dev(c_loc(ff), ip11*ip60, ip11, 1, 5, b, 1)
    dtq = codens_tq_in(tq, ip12, b)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dps = codens_dev(c_loc(ps), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_sch_roe(b, codens_phys_now(), codens_scheme_now(), int(ityprk, c_int32_t), du, dv, &
         int(ip11, c_int64_t), dff, dtq, dsn, dps, c_null_ptr), 'sch_roe')
    call codens_back(c_loc(u), ip11, 1, 5, b)
    if (ityprk .ne. 0) call codens_back(c_loc(ff), ip11, 1, 5, b)
  end subroutine sch_roe
end module mod_sch_roe

module mod_sch_expli
  implicit none
contains
  ! sch_expli.f90:4 -- explicit update v += dt/vol*u of block lm.
  subroutine sch_expli(lm, u, v, dt, vol)
    use para_var
    use codens_c
    implicit none
    integer :: lm
    double precision, target :: u(ip11,ip60), v(ip11,ip60), dt(ip11), vol(ip11)
    type(codens_block) :: b
    type(c_ptr) :: du, dv, ddt, dvol
    b = codens_block_of(lm)
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 1, 5, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
    ddt = codens_dev(c_loc(dt), ip11, ip11, 1, 1, b, 1)
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    call codens_check(codens_sch_expli(b, du, dv, int(ip11, c_int64_t), ddt, dvol, c_null_ptr), 'sch_expli')
    call codens_back(c_loc(v), ip11, 1, 5, b)
  end subroutine sch_expli
end module mod_sch_expli

module mod_implimf_3d
  implicit none
contains
  ! implimf_3d.f90:4 -- matrix-free point-Jacobi implicit phase of block lm
  ! (lmx sub-iterations), then v += d.
  subroutine implimf_3d(lm, u, dt, v, d, ff, mu, mut, equat, lmx, sn, lgsnlt, vol, dtpas, ityprk, &
       dfxx, dfyy, dfzz, dfxy, dfxz, dfyz, dfex, dfey, dfez, rv, ps, cson)
    use para_var
    use para_fige
    use schemanum, only: kfmg
    use codens_c
    implicit none
    integer :: lm, lmx, lgsnlt, ityprk
    double precision :: dtpas
    double precision, target :: u(ip11,ip60), dt(ip11), v(ip11,ip60), d(ip11,ip60), ff(ip11,ip60)
    double precision, target :: mu(ip12), mut(ip12), sn(lgsnlt,nind,ndir), vol(ip11), ps(ip11), cson(ip11)
    double precision :: dfxx(ip00), dfyy(ip00), dfzz(ip00), dfxy(ip00), dfxz(ip00), dfyz(ip00)
    double precision :: dfex(ip00), dfey(ip00), dfez(ip00), rv(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: du, ddt, dv, dd, dff, dmu, dmut, dsn, dvol, dps, dc
    call codens_need_3d(equat, 'implimf_3d')
    b = codens_block_of(lm)
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 1, 5, b, 1)
    ddt = codens_dev(c_loc(dt), ip11, ip11, 1, 1, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
    dd = codens_dev(c_loc(d), ip11*ip60, ip11, 1, 5, b, 1)
    dff = c_null_ptr
    if (ityprk .ne. 0) dff = codens_dev(c_loc(ff), ip11*ip60, ip11, 1, 5, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dps = codens_dev(c_loc(ps), ip11, ip11, 1, 1, b, 1)
    dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_implimf_3d(b, codens_phys_now(), du, ddt, dv, int(ip11, c_int64_t), dd, dff, dmu, &
         dmut, int(lmx, c_int32_t), dsn, dvol, dtpas, int(kfmg, c_int32_t), int(ityprk, c_int32_t), dps, dc, &
         c_null_ptr), 'implimf_3d')
    call codens_back(c_loc(v), ip11, 1, 5, b)
    call codens_back(c_loc(d), ip11, 1, 5, b)
  end subroutine implimf_3d
end module mod_implimf_3d

module mod_implimf_eu
  implicit none
contains
  ! implimf_eu.f90:6 -- Euler point-Jacobi implicit phase of block lm (lmx
  ! sub-iterations), then v += d on all ip60 planes as the reference does.
  subroutine implimf_eu(lm, u, dt, v, d, ff, equat, lmx, sn, lgsnlt, vol, dtpas, ityprk, &
       dfxx, dfyy, dfzz, dfxy, dfxz, dfyz, dfex, dfey, dfez, coefdiag, ps, cson)
    use para_var
    use para_fige
    use schemanum, only: kfmg, dt1min
    use codens_c
    implicit none
    integer :: lm, lmx, lgsnlt, ityprk
    double precision :: dtpas
    double precision, target :: u(ip11,ip60), dt(ip11), v(ip11,ip60), d(ip11,ip60), ff(ip11,ip60)
    double precision, target :: sn(lgsnlt,nind,ndir), vol(ip11), ps(ip11), cson(ip11)
    double precision :: dfxx(ip00), dfyy(ip00), dfzz(ip00), dfxy(ip00), dfxz(ip00), dfyz(ip00)
    double precision :: dfex(ip00), dfey(ip00), dfez(ip00), coefdiag(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: du, ddt, dv, dd, dff, dsn, dvol, dps, dc
    call codens_need_3d(equat, 'implimf_eu')
    b = codens_block_of(lm)
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 1, 5, b, 1)
    ddt = codens_dev(c_loc(dt), ip11, ip11, 1, 1, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, ip60, b, 1)
    dd = codens_dev(c_loc(d), ip11*ip60, ip11, 1, ip60, b, 1)
    dff = c_null_ptr
    if (ityprk .ne. 0) dff = codens_dev(c_loc(ff), ip11*ip60, ip11, 1, 5, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dps = codens_dev(c_loc(ps), ip11, ip11, 1, 1, b, 1)
    dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_implimf_eu(b, codens_phys_now(), du, ddt, dv, int(ip11, c_int64_t), &
         int(ip60, c_int32_t), dd, dff, int(lmx, c_int32_t), dsn, dvol, dt1min, int(kfmg, c_int32_t), &
         int(ityprk, c_int32_t), dps, dc, c_null_ptr), 'implimf_eu')
    call codens_back(c_loc(v), ip11, 1, ip60, b)
    call codens_back(c_loc(d), ip11, 1, 5, b)
  end subroutine implimf_eu
end module mod_implimf_eu

module mod_dissip_jameson
  implicit none
contains
  ! dissip_jameson.f90:4 -- JST dissipation d(:,1:5) of block lm with the
  ! block's coefficients ki2(lm), ki4(lm).
  subroutine dissip_jameson(lm, v, d, equat, sn, lgsnlt, snu, ps, cson)
    use para_var
    use para_fige
    use schemanum, only: ki2, ki4
    use codens_c
    implicit none
    integer :: lm, lgsnlt
    double precision, target :: v(ip11,ip60), d(ip11,ip60), sn(lgsnlt,nind,ndir), ps(ip11), cson(ip11)
    double precision :: snu(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: dv, dd, dsn, dps, dc
    call codens_need_3d(equat, 'dissip_jameson')
    b = codens_block_of(lm)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
    dd = codens_dev(c_loc(d), ip11*ip60, ip11, 1, 5, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dps = codens_dev(c_loc(ps), ip11, ip11, 1, 1, b, 1)
    dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_dissip_jameson(b, ki2(lm), ki4(lm), dv, int(ip11, c_int64_t), dd, dsn, dps, dc, &
         c_null_ptr), 'dissip_jameson')
    call codens_back(c_loc(d), ip11, 1, 5, b)
  end subroutine dissip_jameson
end module mod_dissip_jameson

module mod_met_roe2o
  implicit none
contains
  ! met_roe2o.f90:4 -- second-order upwind dissipation d(:,6:7) of the two
  ! transport equations of block l.  (The reference declares it pure; a
  ! routine that calls the device cannot be, and no caller needs it.)
  subroutine met_roe2o(l, t, d, sn, lgsnlt, vol, del6, del7)
    use para_var
    use para_fige
    use codens_c
    implicit none
    integer, intent(in) :: l, lgsnlt
    double precision, target, intent(inout) :: d(ip11,ip60)
    double precision, intent(inout) :: del6(ip00), del7(ip00)
    double precision, target, intent(in) :: t(ip11,ip60), sn(lgsnlt,nind,ndir)
    double precision, intent(in) :: vol(ip11)
    type(codens_block) :: b
    type(c_ptr) :: dt7, dd, dsn
    b = codens_block_of(l)
    dt7 = codens_dev(c_loc(t), ip11*ip60, ip11, 1, 7, b, 1)
    dd = codens_dev(c_loc(d), ip11*ip60, ip11, 6, 2, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    call codens_check(codens_met_roe2o(b, dt7, int(ip11, c_int64_t), dd, dsn, c_null_ptr), 'met_roe2o')
    call codens_back(c_loc(d), ip11, 6, 2, b)
  end subroutine met_roe2o
end module mod_met_roe2o

module mod_met_roe
  implicit none
contains
  ! met_roe.f90:4 -- first-order upwind dissipation d(:,6:7) of block l,
  ! including the reference's one-entry index shift.
  subroutine met_roe(l, t, d, equat, sn, lgsnlt, vol)
    use para_var
    use para_fige
    use codens_c
    implicit none
    integer :: l, lgsnlt
    double precision, target :: t(ip11,ip60), d(ip11,ip60), sn(lgsnlt,nind,ndir)
    double precision :: vol(ip11)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: dt7, dd, dsn
    call codens_need_3d(equat, 'met_roe')
    b = codens_block_of(l)
    dt7 = codens_dev(c_loc(t), ip11*ip60, ip11, 1, 7, b, 1)
    dd = codens_dev(c_loc(d), ip11*ip60, ip11, 6, 2, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    call codens_check(codens_met_roe(b, dt7, int(ip11, c_int64_t), dd, dsn, c_null_ptr), 'met_roe')
    call codens_back(c_loc(d), ip11, 6, 2, b)
  end subroutine met_roe
end module mod_met_roe

module mod_rbord
  use iso_c_binding, only: c_int64_t, c_double
  implicit none
  ! ghost lists (0-based storage indices) and unit normals, kept with their
  ! device copies while they come out the same: coincident (ghost, adjacent
  ! cell, cell across), walls, far fields and symmetry planes (ghost,
  ! adjacent cell); normals as one (3, faces) block per frontier
  integer(c_int64_t), allocatable, target, save :: rb_cg(:), rb_ci(:), rb_cr(:), rb_wg(:), rb_wi(:)
  integer(c_int64_t), allocatable, target, save :: rb_fg(:), rb_fi(:), rb_yg(:), rb_yi(:)
  real(c_double), allocatable, target, save :: rb_fn(:), rb_yn(:)
contains
  ! rbord.f90:4 -- the ghost cells of v, p, T, c on every frontier of grid
  ! level img, on the device: rbvc's averages over the coincident frontiers
  ! nfbc, the frontiers of nba by kind ('para': clpara; 'nrd': utnrd, rbve,
  ! clnrd; 'sym': clsym; 'rc': nothing), then rbvc again (rbord.f90:137-429);
  ! p, T, c of the symmetry ghosts keep their values, as clsym leaves them.
  ! Each kind writes its own ghost cells from real cells, so the kinds run
  ! grouped, one launch per kind (per far-field frontier: each has its state,
  ! from the reference's utnrd on the host).  Other kinds, preconditioned
  ! far fields (kprec >= 1), a far-field state that varies along the frontier
  ! (a rotating frame) and rotated or remote coincident partners stop: they
  ! have no device form.
  subroutine rbord(kpst, img, mnc, ncin, mnr, xnr, ynr, znr, nxn, nyn, nzn, ncbd, sn, vol, v, dt, bceqt, &
       rpi, rti, d0x, d0y, d0z, x, y, z, pres, tp, rod, roud, rovd, rowd, roed, icyc, tm1, tm2, tm3, tm4, &
       tm5, tm6, tm7, tm8, tm9, tm10, tm11, tm12, tm13, pression, ztemp, cson)
    use para_var
    use para_fige
    use maillage
    use boundary
    use schemanum
    use definition
    use mod_mpi
    use mod_utnrd
    use codens_c
    implicit none
    integer :: icyc, img, kpst, mnc(ip43), mnr(ip44), ncbd(ip41), ncin(ip41)
    double precision :: bceqt(ip41,neqt), cson(ip11), d0x(ip40), d0y(ip40), d0z(ip40), dt(ip11)
    double precision :: nxn(ip42), nyn(ip42), nzn(ip42), pres(ip40), pression(ip11), rod(ip40), roed(ip40)
    double precision :: roud(ip40), rovd(ip40), rowd(ip40), rpi(ip40), rti(ip40), sn(ip31*ndir), tm1(ip40)
    double precision :: tm10(ip40), tm11(ip40), tm12(ip40), tm13(ip40), tm2(ip40), tm3(ip40), tm4(ip40)
    double precision :: tm5(ip40), tm6(ip40), tm7(ip40), tm8(ip40), tm9(ip40), tp(ip40), v(ip11,ip60)
    double precision :: vol(ip11), x(ip21), xnr(ip44), y(ip21), ynr(ip44), z(ip21), znr(ip44), ztemp(ip11)
    target :: v, pression, ztemp, cson
    integer(c_int64_t), allocatable :: cg(:), ci(:), cr(:), wg(:), wi(:), fg(:), fi(:), yg(:), yi(:)
    real(c_double), allocatable :: fn(:), yn(:), w0(:, :)
    integer, allocatable :: nff(:)
    integer :: no, mfb, mfbm, l, lm, m, q, mf2, part, nc, nw, nf, ny, kf, nfr
    type(c_ptr) :: dv, dp, dtm, dc
    type(codens_block) :: b
    integer(c_int64_t) :: ld
    ! lists of this call
    nc = 0; nw = 0; nf = 0; ny = 0; nfr = 0
    do q = 1, mtcx
       nc = nc + mmb(nfbc(q))
    enddo
    do no = 1, mtbx
       mfb = nba(no)
       mfbm = mfb + (img-1)*mtb
       select case (kind_of(cl(mfb)))
       case (1)
          nw = nw + mmb(mfbm)
       case (2)
          if (kprec .ne. 0) error stop 'rbord: preconditioned far field (kprec >= 1) has no device form'
          nf = nf + mmb(mfbm); nfr = nfr + 1
       case (3)
          ny = ny + mmb(mfbm)
       case (4)
       case default
          write(*, '("rbord: frontier kind ",a," has no device form")') cl(mfb)
          error stop 'rbord: frontier kind without a device form'
       end select
    enddo
    allocate(cg(nc), ci(nc), cr(nc), wg(nw), wi(nw), fg(nf), fi(nf), yg(ny), yi(ny), fn(3*nf), yn(3*ny))
    allocate(w0(5, nfr), nff(nfr))
    ! coincident (rbvc.f90): face m of frontier f takes the partner's
    ! adjacent cell of face m
    nc = 0
    do q = 1, mtcx
       mfbm = nfbc(q)
       part = 0
       do mf2 = 1, mtcx
          if (bcl_to_bcg(nfbc(mf2)) == ndcc(mfbm)) part = nfbc(mf2)
       enddo
       if (part == 0 .or. mper(mfbm) /= 0 .or. bcg_to_proc(bcl_to_bcg(mfbm)) /= bcg_to_proc(ndcc(mfbm))) &
            error stop 'rbord: a coincident frontier without a device form (rotated, or its partner remote)'
       do m = 1, mmb(mfbm)
          nc = nc + 1
          cg(nc) = ncbd(mpb(mfbm) + m) - 1
          ci(nc) = ncin(mpb(mfbm) + m) - 1
          cr(nc) = ncin(mpb(part) + m) - 1
       enddo
    enddo
    nw = 0; nf = 0; ny = 0; kf = 0
    do no = 1, mtbx
       mfb = nba(no)
       mfbm = mfb + (img-1)*mtb
       l = ndlb(mfb)
       lm = l + (img-1)*lz
       select case (kind_of(cl(mfb)))
       case (1)
          do m = 1, mmb(mfbm)
             nw = nw + 1
             wg(nw) = ncbd(mpb(mfbm) + m) - 1
             wi(nw) = ncin(mpb(mfbm) + m) - 1
          enddo
       case (2)
          call utnrd(bceqt, mfbm, lm, rod, roud, rovd, rowd, roed, ncbd, y, z)
          kf = kf + 1
          nff(kf) = mmb(mfbm)
          w0(:, kf) = [rod(1), roud(1), rovd(1), rowd(1), roed(1)]
          do m = 1, mmb(mfbm)
             if (rod(m) /= rod(1) .or. roud(m) /= roud(1) .or. rovd(m) /= rovd(1) .or. rowd(m) /= rowd(1) .or. &
                 roed(m) /= roed(1)) error stop 'rbord: a far-field state varying along the frontier (omg /= 0)'
             fg(nf + m) = ncbd(mpb(mfbm) + m) - 1
             fi(nf + m) = ncin(mpb(mfbm) + m) - 1
             fn(3*nf + m) = nxn(mpn(mfbm) + m)
             fn(3*nf + mmb(mfbm) + m) = nyn(mpn(mfbm) + m)
             fn(3*nf + 2*mmb(mfbm) + m) = nzn(mpn(mfbm) + m)
          enddo
          nf = nf + mmb(mfbm)
       case (3)
          do m = 1, mmb(mfbm)
             yg(ny + m) = ncbd(mpb(mfbm) + m) - 1
             yi(ny + m) = ncin(mpb(mfbm) + m) - 1
          enddo
          ! clsym takes the normals of each face; one (3, faces) block for all
          ny = ny + mmb(mfbm)
       end select
    enddo
    ny = 0
    do no = 1, mtbx
       mfb = nba(no)
       mfbm = mfb + (img-1)*mtb
       if (kind_of(cl(mfb)) /= 3) cycle
       do m = 1, mmb(mfbm)
          yn(ny + m) = nxn(mpn(mfbm) + m)
          yn(size(yg) + ny + m) = nyn(mpn(mfbm) + m)
          yn(2*size(yg) + ny + m) = nzn(mpn(mfbm) + m)
       enddo
       ny = ny + mmb(mfbm)
    enddo
    call keep_i(rb_cg, cg); call keep_i(rb_ci, ci); call keep_i(rb_cr, cr)
    call keep_i(rb_wg, wg); call keep_i(rb_wi, wi); call keep_i(rb_fg, fg); call keep_i(rb_fi, fi)
    call keep_i(rb_yg, yg); call keep_i(rb_yi, yi); call keep_d(rb_fn, fn); call keep_d(rb_yn, yn)
    ! every block's window of v, p, T, c on the device (frontiers reach
    ! across blocks); the last call gives the buffers' base addresses
    do l = 1, lzx
       b = codens_block_of(l + (img-1)*lz)
       dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 5, b, 1)
       dp = codens_dev(c_loc(pression), ip11, ip11, 1, 1, b, 1)
       dtm = codens_dev(c_loc(ztemp), ip11, ip11, 1, 1, b, 1)
       dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    enddo
    ld = int(ip11, c_int64_t)
    if (size(rb_cg) > 0) call codens_check(codens_bc_coincident(int(size(rb_cg), c_int64_t), ldev(rb_cg), &
         ldev(rb_ci), ldev(rb_cr), dv, ld, dp, dtm, dc, c_null_ptr), 'rbord')
    if (size(rb_wg) > 0) call codens_check(codens_bc_wall(int(size(rb_wg), c_int64_t), ldev(rb_wg), ldev(rb_wi), &
         dv, ld, dp, dtm, dc, c_null_ptr), 'rbord')
    nf = 0
    do kf = 1, nfr
       call codens_check(codens_bc_farfield(int(nff(kf), c_int64_t), codens_ptr_off(ldev(rb_fg), 8_c_int64_t*nf), &
            codens_ptr_off(ldev(rb_fi), 8_c_int64_t*nf), codens_ptr_off(ddev(rb_fn), 24_c_int64_t*nf), w0(:, kf), &
            codens_phys_now(), dv, ld, dp, dtm, dc, c_null_ptr), 'rbord')
       nf = nf + nff(kf)
    enddo
    if (size(rb_yg) > 0) call codens_check(codens_bc_sym(int(size(rb_yg), c_int64_t), ldev(rb_yg), ldev(rb_yi), &
         ddev(rb_yn), dv, ld, c_null_ptr), 'rbord')
    if (size(rb_cg) > 0) call codens_check(codens_bc_coincident(int(size(rb_cg), c_int64_t), ldev(rb_cg), &
         ldev(rb_ci), ldev(rb_cr), dv, ld, dp, dtm, dc, c_null_ptr), 'rbord')
    do l = 1, lzx
       b = codens_block_of(l + (img-1)*lz)
       call codens_back(c_loc(v), ip11, 1, 5, b)
       call codens_back(c_loc(pression), ip11, 1, 1, b)
       call codens_back(c_loc(ztemp), ip11, 1, 1, b)
       call codens_back(c_loc(cson), ip11, 1, 1, b)
    enddo
  end subroutine rbord

  ! rbord.f90:139-280's tests of a frontier's kind: 1 'para', 2 'nrd*',
  ! 3 'sym*', 4 'rc*', 0 any other
  integer function kind_of(c)
    character(len=4), intent(in) :: c
    kind_of = 0
    if (c .eq. 'para') then
       kind_of = 1
    else if (c(1:3) .eq. 'nrd') then
       kind_of = 2
    else if (c(1:3) .eq. 'sym') then
       kind_of = 3
    else if (c(1:3) .ne. 'rec' .and. c(1:3) .ne. 'rnc' .and. c(1:2) .eq. 'rc') then
       kind_of = 4
    endif
  end function kind_of

  ! device addresses of the kept lists (copied when new or changed)
  function ldev(a) result(dev)
    use codens_c
    integer(c_int64_t), target, intent(in) :: a(:)
    type(c_ptr) :: dev
    integer(c_int64_t) :: nb
    nb = 8_c_int64_t * size(a)
    call codens_check(codens_host_mirror(c_loc(a), nb, 0_c_int64_t, nb, nb, 1_c_int32_t, 1_c_int32_t, dev, &
         c_null_ptr), 'rbord')
  end function ldev

  function ddev(a) result(dev)
    use codens_c
    real(c_double), target, intent(in) :: a(:)
    type(c_ptr) :: dev
    integer(c_int64_t) :: nb
    nb = 8_c_int64_t * size(a)
    call codens_check(codens_host_mirror(c_loc(a), nb, 0_c_int64_t, nb, nb, 1_c_int32_t, 1_c_int32_t, dev, &
         c_null_ptr), 'rbord')
  end function ddev

  ! keep: the new list in place of the kept one unless they are equal (a
  ! replaced list's device copy is forgotten first)
  subroutine keep_i(kept, new)
    use codens_c
    integer(c_int64_t), allocatable, target, intent(inout) :: kept(:)
    integer(c_int64_t), allocatable, intent(inout) :: new(:)
    if (allocated(kept)) then
       if (size(kept) == size(new)) then
          if (all(kept == new)) return
       endif
       if (size(kept) > 0) call codens_modified(c_loc(kept))
       deallocate(kept)
    endif
    call move_alloc(new, kept)
  end subroutine keep_i

  subroutine keep_d(kept, new)
    use codens_c
    real(c_double), allocatable, target, intent(inout) :: kept(:)
    real(c_double), allocatable, intent(inout) :: new(:)
    if (allocated(kept)) then
       if (size(kept) == size(new)) then
          if (all(kept == new)) return
       endif
       if (size(kept) > 0) call codens_modified(c_loc(kept))
       deallocate(kept)
    endif
    call move_alloc(new, kept)
  end subroutine keep_d
end module mod_rbord
