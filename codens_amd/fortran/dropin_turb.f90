! Drop-in replacements, on the GPU, for the turbulence-stage routines of
! met_num (same module names, subroutine names and dummy-argument lists as
! the reference; see dropin_flow.f90 for the transfer rules).
!
! Conventions of the reference these routines keep:
! * the velocity gradients dvxx..dvzz and tprod are block-local arrays
!   (index m = n - npc); here they travel as one 9-slot device bundle of
!   stride nnn, the C-ABI's g9 layout;
! * teq_gradv, met_gradtr and met_smsa receive the whole sn(ip31*ndir) and
!   use the block's slice from npsn = ndir*npfb(l)+1;
! * the gradient weights cmui1..cmuk2 are 1 in the reference (inivec.f90:110;
!   metric2 is never called), so the kernels do not read them.

module mod_teq_gradv
  implicit none
contains
  ! teq_gradv.f90:4 -- gradients of u, v, w (and the rho >= 1e-20 clamp)
  subroutine teq_gradv(l, sn, vol, t, s, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), t(ip11,ip60), s(ip00)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dt, dg, g(9)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dt = codens_dev(c_loc(t), ip11*ip60, ip11, 1, 4, b, 1)
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    call codens_check(codens_teq_gradv(b, dsn, dvol, dt, int(ip11, c_int64_t), dg, c_null_ptr), 'teq_gradv')
    call codens_back(c_loc(t), ip11, 1, 1, b)
    call codens_bundle_back(g, 9, nnn(l), 0, nnn(l))
  end subroutine teq_gradv
end module mod_teq_gradv

module mod_met_prod
  implicit none
contains
  ! met_prod.f90:4 -- production of k
  subroutine met_prod(l, ncyc, mut, s, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, tprod)
    use para_var
    use maillage
    use codens_c
    implicit none
    integer :: l, ncyc
    double precision, target :: mut(ip12), s(ip11,ip60), tprod(ip00)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    type(codens_block) :: b
    type(c_ptr) :: dmut, dv, dg, dtp, g(9)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    dv = codens_dev(c_loc(s), ip11*ip60, ip11, 1, 7, b, 1)
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    call codens_check(codens_met_prod(b, dmut, dv, int(ip11, c_int64_t), dg, dtp, c_null_ptr), 'met_prod')
    call codens_check(codens_host_fetch(c_loc(tprod), 0_c_int64_t, 8_c_int64_t*nnn(l), 8_c_int64_t*nnn(l), &
         1_c_int32_t, c_null_ptr), 'met_prod')
  end subroutine met_prod
end module mod_met_prod

module mod_met_gradtr
  implicit none
contains
  ! met_gradtr.f90:4 -- gradients of the two transported variables
  subroutine met_gradtr(l, sn, vol, s, mu, mut, tt, dtdx, dtdy, dtdz, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), s(ip11,ip60), mu(ip12), mut(ip12)
    double precision :: tt(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00)
    double precision, target :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dfd, f(6)
    b = codens_block_of(l)
    f = [c_loc(fd5x), c_loc(fd5y), c_loc(fd5z), c_loc(fd6x), c_loc(fd6y), c_loc(fd6z)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(s), ip11*ip60, ip11, 1, 7, b, 1)
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    call codens_check(codens_met_gradtr(b, dsn, dvol, dv, int(ip11, c_int64_t), dfd, c_null_ptr), 'met_gradtr')
    call codens_bundle_back(f, 6, ip12, npc(l), npc(l) + nnn(l))
  end subroutine met_gradtr
end module mod_met_gradtr

module mod_met_smsa
  implicit none
contains
  ! met_smsa.f90:4 -- Spalart-Allmaras source and implicit coefficient.
  ! Its cross-diffusion term reads txxf5x..txzf5z at the block-local index
  ! (a quirk kept by the kernel), so their window starts at entry 0.
  subroutine met_smsa(l, sn, vol, v, mu, dist, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, &
       txxf5x, txyf5y, txzf5z, cfke, t, dtdx, dtdy, dtdz, sdif, qcxts5, qcyts6, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), dist(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), cfke(ip13), qcxts5(ip12), qcyts6(ip12)
    double precision :: t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00), sdif(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, ddist, dg, dfd, dcf, d5, d6, g(9), f(3)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dfd = codens_bundle_in(f, 3, ip12, 0, npc(l) + nnn(l))
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smsa(b, codens_tc_now(), dsn, dvol, dv, int(ip11, c_int64_t), dmu, ddist, dg, dfd, &
         dcf, d5, d6, c_null_ptr), 'met_smsa')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smsa
end module mod_met_smsa

module mod_met_smko
  implicit none
contains
  ! met_smko.f90:4 -- compressible Wilcox k-omega sources
  subroutine met_smko(l, v, txxf5x, txyf5y, txzf5z, tyyf6x, tyzf6y, tzzf6z, tprod, cfke, qcxts5, qcyts6, cson)
    use para_var
    use maillage
    use codens_c
    implicit none
    integer :: l
    double precision, target :: v(ip11,ip60), tprod(ip00), cfke(ip13), qcxts5(ip12), qcyts6(ip12), cson(ip11)
    double precision :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), tyyf6x(ip12), tyzf6y(ip12), tzzf6z(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dv, dtp, dcf, d5, d6, dc
    b = codens_block_of(l)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    dc = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    call codens_check(codens_met_smko(b, codens_tc_now(), dv, int(ip11, c_int64_t), dtp, dcf, d5, d6, dc, &
         c_null_ptr), 'met_smko')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smko
end module mod_met_smko

module mod_met_smkor
  implicit none
contains
  ! met_smkor.f90:4 -- k-omega sources with cross diffusion and strain limiter
  subroutine met_smkor(l, v, txxf5x, txyf5y, txzf5z, tyyf6x, tyzf6y, tzzf6z, &
       dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, tprod, cfke, qcxts5, qcyts6)
    use para_var
    use maillage
    use modeleturb, only: sigmd
    use codens_c
    implicit none
    integer :: l
    double precision, target :: v(ip11,ip60), tprod(ip00), cfke(ip13), qcxts5(ip12), qcyts6(ip12)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), tyyf6x(ip12), tyzf6y(ip12), tzzf6z(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    type(codens_block) :: b
    type(c_ptr) :: dv, dfd, dg, dtp, dcf, d5, d6, g(9), f(6)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z), c_loc(tyyf6x), c_loc(tyzf6y), c_loc(tzzf6z)]
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smkor(b, codens_tc_now(), sigmd, dv, int(ip11, c_int64_t), dfd, dg, dtp, dcf, &
         d5, d6, c_null_ptr), 'met_smkor')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smkor
end module mod_met_smkor

module mod_met_smch
  implicit none
contains
  ! met_smch.f90:4 -- Chien low-Reynolds k-epsilon sources
  subroutine met_smch(l, v, mu, dist, mnpar, utau, tprod, qcxts5, qcyts6)
    use para_var
    use maillage
    use modeleturb, only: cke1, cke2, utaumin
    use codens_c
    implicit none
    integer :: l
    integer, target :: mnpar(ip12)
    double precision, target :: v(ip11,ip60), mu(ip12), dist(ip12), utau(ip42), tprod(ip00)
    double precision, target :: qcxts5(ip12), qcyts6(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dv, dmu, ddist, dmn, dut, dtp, d5, d6
    b = codens_block_of(l)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    call codens_check(codens_host_mirror(c_loc(mnpar), 4_c_int64_t*ip12, 4_c_int64_t*npc(l), 4_c_int64_t*nnn(l), &
         4_c_int64_t*nnn(l), 1_c_int32_t, 1_c_int32_t, dmn, c_null_ptr), 'met_smch')
    dut = codens_dev_all(c_loc(utau), ip42, 1)
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smch(b, cke1, cke2, utaumin, dv, int(ip11, c_int64_t), dmu, ddist, dmn, dut, dtp, &
         d5, d6, c_null_ptr), 'met_smch')
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smch
end module mod_met_smch

module mod_met_fludc
  implicit none
contains
  ! met_fludc.f90:4 (kfludis = 1) -- flux densities with constant sigma_k, sigma_e
  subroutine met_fludc(l, s, mu, mut, t, dtdx, dtdy, dtdz, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z)
    use para_var
    use codens_c
    implicit none
    integer :: l
    double precision :: s(ip11,ip60), mu(ip12), mut(ip12), t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00)
    double precision :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    call codens_flud(1, l, s, mu, mut, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z, 'met_fludc')
  end subroutine met_fludc
end module mod_met_fludc

module mod_met_fludcsa
  implicit none
contains
  ! met_fludcsa.f90:4 (kfludis = 2) -- SA flux density, fd6 = 0
  subroutine met_fludcsa(l, s, mu, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z)
    use para_var
    use codens_c
    implicit none
    integer :: l
    double precision :: s(ip11,ip60), mu(ip12)
    double precision :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    call codens_flud(2, l, s, mu, mu, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z, 'met_fludcsa')
  end subroutine met_fludcsa
end module mod_met_fludcsa

module mod_met_fludko
  implicit none
contains
  ! met_fludko.f90:4 (kfludis = 5) -- Wilcox k-omega flux densities
  subroutine met_fludko(l, mu, mut, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z)
    use para_var
    use codens_c
    implicit none
    integer :: l
    double precision :: mu(ip12), mut(ip12)
    double precision :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    double precision :: nov(1, 1)
    call codens_flud(5, l, nov, mu, mut, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z, 'met_fludko')
  end subroutine met_fludko
end module mod_met_fludko

module mod_sch_turb
  implicit none
contains
  ! sch_turb.f90:4 -- flux balance u(6:7) of the transport equations
  subroutine sch_turb(l, u, v, d, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z, ts6, ts7, equat, ncin, sn, lgsnlt, vol, &
       f6x, f6y, f6z, f7x, f7y, f7z)
    use para_var
    use para_fige
    use maillage
    use codens_c
    implicit none
    integer :: l, lgsnlt
    integer :: ncin(ip41)
    double precision, target :: u(ip11,ip60), v(ip11,ip60), d(ip11,ip60), ts6(ip12), ts7(ip12)
    double precision, target :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    double precision, target :: sn(lgsnlt,nind,ndir), vol(ip11)
    double precision :: f6x(ip00), f6y(ip00), f6z(ip00), f7x(ip00), f7y(ip00), f7z(ip00)
    character(len=7) :: equat
    type(codens_block) :: b
    type(c_ptr) :: du, dv, dd, dfd, d6, d7, dsn, dvol, f(6)
    call codens_need_3d(equat, 'sch_turb')
    b = codens_block_of(l)
    f = [c_loc(fd5x), c_loc(fd5y), c_loc(fd5z), c_loc(fd6x), c_loc(fd6y), c_loc(fd6z)]
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 6, 2, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dd = codens_dev(c_loc(d), ip11*ip60, ip11, 6, 2, b, 1)
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    d6 = codens_dev(c_loc(ts6), ip12, ip12, 1, 1, b, 1)
    d7 = codens_dev(c_loc(ts7), ip12, ip12, 1, 1, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    call codens_check(codens_sch_turb(b, du, dv, int(ip11, c_int64_t), dd, dfd, d6, d7, dsn, dvol, c_null_ptr), &
         'sch_turb')
    call codens_back(c_loc(u), ip11, 6, 2, b)
  end subroutine sch_turb
end module mod_sch_turb

module mod_impli2_eqt_3d
  implicit none
contains
  ! impli2_eqt_3d.f90:4 -- line-implicit phase of the transport equations
  ! (the wall-law zeroing for lparoi = 1, 2 is not provided: stops)
  subroutine impli2_eqt_3d(l, u, dt, v, mu, mut, cfke, ncin, ncyc, sn, lgsnlt, vol, &
       dwi6, dwi7, u1, u2, u3, rv, coefdiag, alpha, beta6, beta7)
    use para_var
    use para_fige
    use maillage
    use modeleturb, only: lparoi
    use chainecarac, only: equatt
    use codens_c
    implicit none
    integer :: l, ncyc, lgsnlt
    integer :: ncin(ip41)
    double precision, target :: u(ip11,ip60), dt(ip11), v(ip11,ip60), mu(ip12), mut(ip12), cfke(ip13)
    double precision, target :: sn(lgsnlt,nind,ndir), vol(ip11)
    double precision :: dwi6(ip00), dwi7(ip00), u1(ip00), u2(ip00), u3(ip00), rv(ip00), coefdiag(ip00)
    double precision :: alpha(ip00), beta6(ip00), beta7(ip00)
    type(codens_block) :: b
    type(c_ptr) :: du, ddt, dv, dmu, dmut, dcf, dsn, dvol
    double precision :: cmt
    if (lparoi .eq. 1 .or. lparoi .eq. 2) error stop 'impli2_eqt_3d: wall laws are not provided on the GPU'
    select case (equatt(1:3))
    case ('2Sm')
       cmt = 1.43
    case ('2WL', '2KO')
       cmt = 2.
    case ('2MT')
       cmt = 1.17
    case default
       cmt = 1.
    end select
    b = codens_block_of(l)
    du = codens_dev(c_loc(u), ip11*ip60, ip11, 6, 2, b, 1)
    ddt = codens_dev(c_loc(dt), ip11, ip11, 1, 1, b, 1)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    dsn = codens_dev_all(c_loc(sn), 9*lgsnlt, codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    call codens_check(codens_impli2_eqt_3d(b, du, int(ip11, c_int64_t), ddt, dv, dmu, dmut, dcf, dsn, dvol, cmt, &
         c_null_ptr), 'impli2_eqt_3d')
    call codens_back(c_loc(v), ip11, 6, 2, b)
  end subroutine impli2_eqt_3d
end module mod_impli2_eqt_3d

module mod_met_samut
  implicit none
contains
  ! met_samut.f90:4 -- SA eddy viscosity
  subroutine met_samut(l, v, mu, mut)
    use para_var
    use codens_c
    implicit none
    integer :: l
    double precision, target :: v(ip11,ip60), mu(ip12), mut(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dv, dmu, dmut
    b = codens_block_of(l)
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_samut(b, codens_tc_now(), dv, int(ip11, c_int64_t), dmu, dmut, c_null_ptr), &
         'met_samut')
    call codens_back(c_loc(mut), ip12, 1, 1, b)
  end subroutine met_samut
end module mod_met_samut

module mod_met_kocmut
  implicit none
contains
  ! met_kocmut.f90:4 -- k-omega eddy viscosity (SST limiter for kwsst = 1,
  ! whose velocity gradients land in dvxx..dvzz as in the reference)
  subroutine met_kocmut(l, sn, vol, t, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, dist, v, mu, mut, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use modeleturb, only: kwsst
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), t(ip00), dist(ip12), v(ip11,ip60), mu(ip12), mut(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, dmut, ddist, dg, g(9)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    call codens_check(codens_met_kocmut(b, codens_tc_now(), int(kwsst, c_int32_t), dsn, dvol, dv, &
         int(ip11, c_int64_t), dmu, dmut, ddist, dg, c_null_ptr), 'met_kocmut')
    call codens_back(c_loc(mut), ip12, 1, 1, b)
    if (kwsst .eq. 1) then
       call codens_back(c_loc(v), ip11, 1, 1, b)
       call codens_bundle_back(g, 9, nnn(l), 0, nnn(l))
    endif
  end subroutine met_kocmut
end module mod_met_kocmut

module mod_met_cutsa
  implicit none
contains
  ! met_cutsa.f90:4 -- SA limiter
  subroutine met_cutsa(l, s)
    use para_var
    use modeleturb, only: epsk, epse
    use codens_c
    implicit none
    integer :: l
    double precision, target :: s(ip11,ip60)
    call codens_cut(2, l, s, epsk, epse)
  end subroutine met_cutsa
end module mod_met_cutsa

module mod_met_cut
  implicit none
contains
  ! met_cut.f90:4 -- two-equation limiter
  subroutine met_cut(l, s)
    use para_var
    use modeleturb, only: epsk, epse
    use codens_c
    implicit none
    integer :: l
    double precision, target :: s(ip11,ip60)
    call codens_cut(1, l, s, epsk, epse)
  end subroutine met_cut
end module mod_met_cut

module mod_met_mtcorf1
  implicit none
contains
  ! met_mtcorf1.f90:4 -- wall correction of the Menter blending function
  ! along the wall-normal line of every wall point (walls of block l; all
  ! walls for l = 0), one wall after another
  subroutine met_mtcorf1(l, ncin, dist, mnpar, frac)
    use para_var
    use modeleturb, only: imxclko
    use codens_c
    implicit none
    integer :: l
    integer :: ncin(ip41)
    integer, target :: mnpar(ip12)
    double precision, target :: dist(ip12), frac(ip12)
    type(c_ptr) :: dfr, dmn, dl
    integer :: nf
    call codens_wall_lines(l, ncin, nf)
    if (nf .eq. 0) return
    dl = codens_lines_dev()
    dfr = codens_dev_all(c_loc(frac), ip12, 1)
    call codens_check(codens_host_mirror(c_loc(mnpar), 4_c_int64_t*ip12, 0_c_int64_t, 4_c_int64_t*ip12, &
         4_c_int64_t*ip12, 1_c_int32_t, 1_c_int32_t, dmn, c_null_ptr), 'met_mtcorf1')
    call codens_check(codens_met_mtcorf1(int(nf, c_int32_t), codens_face_start, dl, int(imxclko, c_int32_t), dmn, &
         dfr, int(ip12, c_int64_t), c_null_ptr), 'met_mtcorf1')
    call codens_check(codens_host_fetch(c_loc(frac), 0_c_int64_t, 8_c_int64_t*ip12, 8_c_int64_t*ip12, 1_c_int32_t, &
         c_null_ptr), 'met_mtcorf1')
  end subroutine met_mtcorf1
end module mod_met_mtcorf1

module mod_met_smmt
  implicit none
contains
  ! met_smmt.f90:4 -- Menter k-omega sources, blending function (with the
  ! wall correction of met_mtcorf1 on the walls of block l) and gkgo
  subroutine met_smmt(l, ncyc, v, mu, mut, dist, mnpar, ncin, txxf5x, txyf5y, txzf5z, tyyf6x, tyzf6y, tzzf6z, &
       tprod, cfke, gkgo, frac, qcxts5, qcyts6)
    use para_var
    use maillage
    use modeleturb, only: kfracom, imxclko
    use codens_c
    implicit none
    integer :: l, ncyc
    integer, target :: mnpar(ip12)
    integer :: ncin(ip41)
    double precision, target :: v(ip11,ip60), mu(ip12), mut(ip12), dist(ip12), tprod(ip00), cfke(ip13)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), tyyf6x(ip12), tyzf6y(ip12), tzzf6z(ip12)
    double precision, target :: gkgo(ip00), frac(ip12), qcxts5(ip12), qcyts6(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dv, dmu, dmut, ddist, dfd, dtp, dcf, dgk, dfr, d5, d6, dmn, dl, f(6)
    integer :: nf
    b = codens_block_of(l)
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z), c_loc(tyyf6x), c_loc(tyzf6y), c_loc(tzzf6z)]
    nf = 0
    dl = c_null_ptr
    dmn = c_null_ptr
    if (kfracom .ne. 0) then
       call codens_wall_lines(l, ncin, nf)
       if (nf .gt. 0) then
          dl = codens_lines_dev()
          call codens_check(codens_host_mirror(c_loc(mnpar), 4_c_int64_t*ip12, 0_c_int64_t, 4_c_int64_t*ip12, &
               4_c_int64_t*ip12, 1_c_int32_t, 1_c_int32_t, dmn, c_null_ptr), 'met_smmt')
       endif
    endif
    if (.not. allocated(codens_face_start)) allocate(codens_lines(1), codens_face_start(1))
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dgk = codens_dev_all(c_loc(gkgo), nnn(l), 1)
    dfr = codens_dev_all(c_loc(frac), ip12, 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smmt(b, codens_menter_now(), int(kfracom, c_int32_t), int(imxclko, c_int32_t), dv, &
         int(ip11, c_int64_t), dmu, dmut, ddist, dfd, dtp, dcf, dgk, dfr, d5, d6, int(nf, c_int32_t), &
         codens_face_start, dl, dmn, c_null_ptr), 'met_smmt')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
    call codens_check(codens_host_fetch(c_loc(gkgo), 0_c_int64_t, 8_c_int64_t*nnn(l), 8_c_int64_t*nnn(l), &
         1_c_int32_t, c_null_ptr), 'met_smmt')
    call codens_check(codens_host_fetch(c_loc(frac), 0_c_int64_t, 8_c_int64_t*ip12, 8_c_int64_t*ip12, 1_c_int32_t, &
         c_null_ptr), 'met_smmt')
  end subroutine met_smmt
end module mod_met_smmt

module mod_met_fludmt
  implicit none
contains
  ! met_fludmt.f90:4 (kfludis = 4) -- Menter flux densities, sigmas blended
  ! with frac when kfracom != 0
  subroutine met_fludmt(l, s, mu, mut, frac, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z)
    use para_var
    use maillage
    use modeleturb, only: kfracom
    use codens_c
    implicit none
    integer :: l
    double precision :: s(ip11,ip60)
    double precision, target :: mu(ip12), mut(ip12), frac(ip12)
    double precision, target :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dmu, dmut, dfr, dfd, f(6)
    b = codens_block_of(l)
    f = [c_loc(fd5x), c_loc(fd5y), c_loc(fd5z), c_loc(fd6x), c_loc(fd6y), c_loc(fd6z)]
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    dfr = c_null_ptr
    if (kfracom .ne. 0) dfr = codens_dev(c_loc(frac), ip12, ip12, 1, 1, b, 1)
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    call codens_check(codens_met_fludmt(b, codens_menter_now(), int(kfracom, c_int32_t), dmu, dmut, dfr, dfd, &
         int(ip12, c_int64_t), c_null_ptr), 'met_fludmt')
    call codens_bundle_back(f, 6, ip12, npc(l), npc(l) + nnn(l))
  end subroutine met_fludmt
end module mod_met_fludmt

module mod_met_komut
  implicit none
contains
  ! met_komut.f90:4 -- Menter eddy viscosity (SST limiter for komsst = 1,
  ! whose velocity gradients land in dvxx..dvzz as in the reference)
  subroutine met_komut(l, sn, vol, t, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, dist, v, mu, mut, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use modeleturb, only: komsst, betae
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), t(ip00), dist(ip12), v(ip11,ip60), mu(ip12), mut(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, dmut, ddist, dg, g(9)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    call codens_check(codens_met_komut(b, betae, int(komsst, c_int32_t), dsn, dvol, dv, int(ip11, c_int64_t), dmu, &
         dmut, ddist, dg, c_null_ptr), 'met_komut')
    call codens_back(c_loc(mut), ip12, 1, 1, b)
    if (komsst .eq. 1) then
       call codens_back(c_loc(v), ip11, 1, 1, b)
       call codens_bundle_back(g, 9, nnn(l), 0, nnn(l))
    endif
  end subroutine met_komut
end module mod_met_komut

module mod_met_smmtr
  implicit none
contains
  ! met_smmtr.f90:4 -- Menter sources with Durbin's realizability limit (the
  ! blending function with the wall correction on the walls of block l)
  subroutine met_smmtr(l, v, mu, mut, dist, mnpar, ncin, txxf5x, txyf5y, txzf5z, tyyf6x, tyzf6y, tzzf6z, &
       dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, tprod, cfke, frac, qcxts5, qcyts6)
    use para_var
    use maillage
    use modeleturb, only: imxclko
    use codens_c
    implicit none
    integer :: l
    integer, target :: mnpar(ip12)
    integer :: ncin(ip41)
    double precision, target :: v(ip11,ip60), mu(ip12), mut(ip12), dist(ip12), tprod(ip00), cfke(ip13)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), tyyf6x(ip12), tyzf6y(ip12), tzzf6z(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision, target :: frac(ip12), qcxts5(ip12), qcyts6(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dv, dmu, dmut, ddist, dfd, dg, dtp, dcf, dfr, d5, d6, dmn, dl, f(6), g(9)
    integer :: nf
    b = codens_block_of(l)
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z), c_loc(tyyf6x), c_loc(tyzf6y), c_loc(tzzf6z)]
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dl = c_null_ptr
    dmn = c_null_ptr
    call codens_wall_lines(l, ncin, nf)
    if (nf .gt. 0) then
       dl = codens_lines_dev()
       call codens_check(codens_host_mirror(c_loc(mnpar), 4_c_int64_t*ip12, 0_c_int64_t, 4_c_int64_t*ip12, &
            4_c_int64_t*ip12, 1_c_int32_t, 1_c_int32_t, dmn, c_null_ptr), 'met_smmtr')
    endif
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dfr = codens_dev_all(c_loc(frac), ip12, 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smmtr(b, codens_menter_now(), int(imxclko, c_int32_t), dv, int(ip11, c_int64_t), &
         dmu, dmut, ddist, dfd, dg, dtp, dcf, dfr, d5, d6, int(nf, c_int32_t), codens_face_start, dl, dmn, &
         c_null_ptr), 'met_smmtr')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
    call codens_check(codens_host_fetch(c_loc(frac), 0_c_int64_t, 8_c_int64_t*ip12, 8_c_int64_t*ip12, 1_c_int32_t, &
         c_null_ptr), 'met_smmtr')
  end subroutine met_smmtr
end module mod_met_smmtr

module mod_met_smdes
  implicit none
contains
  ! met_smdes.f90:4 -- Spalart-Allmaras DES sources: met_smsa's with the
  ! length min(d, 0.65 delta), delta from met_dist (returned in delta), |fv2|
  subroutine met_smdes(l, x, y, z, delta, sn, vol, v, mu, dist, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, &
       dvzz, txxf5x, txyf5y, txzf5z, cfke, t, dtdx, dtdy, dtdz, sdif, qcxts5, qcyts6, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use codens_c
    implicit none
    integer :: l
    double precision, target :: x(ip21), y(ip21), z(ip21), delta(ip00)
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), dist(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), cfke(ip13), qcxts5(ip12), qcyts6(ip12)
    double precision :: t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00), sdif(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dx, dy, dz, ddl, dsn, dvol, dv, dmu, ddist, dg, dfd, dcf, d5, d6, g(9), f(3)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z)]
    dx = codens_dev(c_loc(x), ip21, ip21, 1, 1, b, codens_geom_mode())
    dy = codens_dev(c_loc(y), ip21, ip21, 1, 1, b, codens_geom_mode())
    dz = codens_dev(c_loc(z), ip21, ip21, 1, 1, b, codens_geom_mode())
    ddl = codens_dev_all(c_loc(delta), nnn(l), 1)
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dfd = codens_bundle_in(f, 3, ip12, 0, npc(l) + nnn(l))
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smdes(b, codens_tc_now(), dx, dy, dz, ddl, dsn, dvol, dv, int(ip11, c_int64_t), &
         dmu, ddist, dg, dfd, dcf, d5, d6, c_null_ptr), 'met_smdes')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
    call codens_check(codens_host_fetch(c_loc(delta), 0_c_int64_t, 8_c_int64_t*nnn(l), 8_c_int64_t*nnn(l), &
         1_c_int32_t, c_null_ptr), 'met_smdes')
  end subroutine met_smdes
end module mod_met_smdes

module mod_met_smsasas
  implicit none
contains
  ! met_smsasas.f90:4 -- Spalart-Allmaras with the SAS destruction (the
  ! scratch arrays t, dtdx..sdif, tn1..tn10 are not written)
  subroutine met_smsasas(l, sn, vol, v, mu, dist, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, &
       txxf5x, txyf5y, txzf5z, cfke, t, dtdx, dtdy, dtdz, sdif, qcxts5, qcyts6, tn1, tn2, tn3, tn10, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use schemanum, only: ischema
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), dist(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), cfke(ip13), qcxts5(ip12), qcyts6(ip12)
    double precision :: t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00), sdif(ip00), tn1(ip00), tn2(ip00), tn3(ip00)
    double precision :: tn10(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, ddist, dg, dfd, dcf, d5, d6, g(9), f(3)
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dfd = codens_bundle_in(f, 3, ip12, 0, npc(l) + nnn(l))
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smsasas(b, codens_tc_now(), int(ischema, c_int32_t), dsn, dvol, dv, &
         int(ip11, c_int64_t), dmu, ddist, dg, dfd, dcf, d5, d6, c_null_ptr), 'met_smsasas')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smsasas
end module mod_met_smsasas

module mod_met_smkl
  implicit none
contains
  ! met_smkl.f90:4 -- Smith k-l sources (with met_bark's term); the scratch
  ! arrays t, dtdx..dtdz, bark are not written
  subroutine met_smkl(l, sn, vol, v, mu, mut, dist, mnpar, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, &
       txxf5x, txyf5y, txzf5z, tyyf6x, tyzf6y, tzzf6z, tprod, cfke, t, dtdx, dtdy, dtdz, bark, qcxts5, qcyts6, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use modeleturb, only: cklb1, ckle2, sigmal, xkappa
    use codens_c
    implicit none
    integer :: l
    integer :: mnpar(ip12)
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), mut(ip12), dist(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00), tprod(ip00)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), tyyf6x(ip12), tyzf6y(ip12), tzzf6z(ip12)
    double precision, target :: cfke(ip13), qcxts5(ip12), qcyts6(ip12)
    double precision :: t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00), bark(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, dmut, ddist, dfd, dg, dtp, dcf, d5, d6, f(6), g(9)
    real(c_double) :: kc(4)
    b = codens_block_of(l)
    kc = [cklb1, ckle2, sigmal, xkappa]
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z), c_loc(tyyf6x), c_loc(tyzf6y), c_loc(tzzf6z)]
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smkl(b, kc, dsn, dvol, dv, int(ip11, c_int64_t), dmu, dmut, ddist, dfd, dg, dtp, &
         dcf, d5, d6, c_null_ptr), 'met_smkl')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smkl
end module mod_met_smkl

module mod_met_klmut
  implicit none
contains
  ! met_klmut.f90:4 -- Smith k-l eddy viscosity
  subroutine met_klmut(l, v, mu, mut, dist)
    use para_var
    use modeleturb, only: cklb1, ckle2, sigmal, xkappa
    use codens_c
    implicit none
    integer :: l
    double precision, target :: v(ip11,ip60), mu(ip12), mut(ip12), dist(ip12)
    type(codens_block) :: b
    type(c_ptr) :: dv, dmu, dmut, ddist
    real(c_double) :: kc(4)
    b = codens_block_of(l)
    kc = [cklb1, ckle2, sigmal, xkappa]
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    call codens_check(codens_met_klmut(b, kc, dv, int(ip11, c_int64_t), dmu, dmut, ddist, c_null_ptr), 'met_klmut')
    call codens_back(c_loc(mut), ip12, 1, 1, b)
  end subroutine met_klmut
end module mod_met_klmut

module mod_met_smke
  implicit none
contains
  ! met_smke.f90:4 -- k-epsilon sources: met_exgr on the velocity gradients
  ! (their ghost layers come back written, as the reference leaves them),
  ! met_bark and met_bare into bark and bare, then the variant of
  ! equatt(4:4) (met_smkes, met_smker, met_smkec, met_smkesas); the scratch
  ! arrays t, dtdx..dtdz and tn1..tn3 are not written
  subroutine met_smke(l, sn, vol, v, mu, mut, tprod, cfke, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, &
       t, dtdx, dtdy, dtdz, bark, bare, qcxts5, qcyts6, tn1, tn2, tn3, cson, cmui1, cmui2, cmuj1, cmuj2, cmuk1, &
       cmuk2)
    use para_var
    use para_fige
    use maillage
    use chainecarac, only: equatt
    use modeleturb, only: cke1, cke2, cmu, lparoi
    use schemanum, only: ischema
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), mut(ip12), tprod(ip00)
    double precision, target :: cfke(ip13), cson(ip11), qcxts5(ip12), qcyts6(ip12), bark(ip00), bare(ip00)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision :: t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00), tn1(ip00), tn2(ip00), tn3(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, dmut, dtp, dg, dcs, dbb, dcf, d5, d6, g(9), bb(2)
    real(c_double) :: kc(3)
    integer(c_int32_t) :: variant
    select case (equatt(4:4))
    case ('R')
       variant = 1
    case ('C')
       variant = 2
    case ('L')
       variant = 3
    case default
       variant = 0
    end select
    b = codens_block_of(l)
    kc = [cke1, cke2, cmu]
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    bb = [c_loc(bark), c_loc(bare)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    dcs = codens_dev(c_loc(cson), ip11, ip11, 1, 1, b, 1)
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dbb = codens_bundle_in(bb, 2, nnn(l), 0, nnn(l))
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smke(b, kc, variant, int(lparoi, c_int32_t), int(ischema, c_int32_t), dsn, dvol, &
         dv, int(ip11, c_int64_t), dmu, dmut, dtp, dg, dcs, dbb, codens_ptr_off(dbb, 8_c_int64_t*nnn(l)), dcf, d5, &
         d6, c_null_ptr), 'met_smke')
    call codens_bundle_back(g, 9, nnn(l), 0, nnn(l))
    call codens_bundle_back(bb, 2, nnn(l), 0, nnn(l))
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smke
end module mod_met_smke

module mod_met_kemut
  implicit none
contains
  ! met_kemut.f90:4 -- k-epsilon eddy viscosity ('2JL' with kesst 0 or 1,
  ! '2LS'); with kesst = 1 the velocity gradients and the rho clamp of
  ! teq_gradv come back written, as the reference leaves them; the scratch t
  ! is not written
  subroutine met_kemut(l, sn, vol, t, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, dvzz, dist, v, mu, mut, &
       cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use chainecarac, only: equatt
    use modeleturb, only: cmu, lparoi, kesst
    use codens_c
    implicit none
    integer :: l
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), mut(ip12), dist(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00)
    double precision :: t(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, dmut, ddist, dg, g(9)
    integer(c_int32_t) :: form
    if (equatt(1:3) .eq. '2LS') then
       form = 2
    else if (equatt(1:3) .eq. '2JL') then
       form = merge(1, 0, kesst .eq. 1)
    else
       return
    end if
    if (equatt(1:3) .eq. '2JL' .and. kesst .ne. 0 .and. kesst .ne. 1) return
    b = codens_block_of(l)
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    call codens_check(codens_met_kemut(b, cmu, form, int(lparoi, c_int32_t), dsn, dvol, dv, int(ip11, c_int64_t), &
         dmu, dmut, ddist, dg, c_null_ptr), 'met_kemut')
    call codens_back(c_loc(mut), ip12, 1, 1, b)
    if (form .eq. 1) then
       call codens_back(c_loc(v), ip11, 1, 1, b)
       call codens_bundle_back(g, 9, nnn(l), 0, nnn(l))
    end if
  end subroutine met_kemut
end module mod_met_kemut

module mod_met_smklsas
  implicit none
contains
  ! met_smklsas.f90:4 -- Smith k-l sources with the scale-adaptive term
  ! (met_bark's term, met_laplaciens of u and v); the scratch arrays t,
  ! dtdx..dtdz, bark and tn1..tn9 are not written
  subroutine met_smklsas(l, sn, vol, v, mu, mut, dist, mnpar, dvxx, dvxy, dvxz, dvyx, dvyy, dvyz, dvzx, dvzy, &
       dvzz, txxf5x, txyf5y, txzf5z, tyyf6x, tyzf6y, tzzf6z, tprod, cfke, t, dtdx, dtdy, dtdz, bark, qcxts5, &
       qcyts6, tn1, tn2, tn3, tn9, cmui1, cmui2, cmuj1, cmuj2, cmuk1, cmuk2)
    use para_var
    use para_fige
    use maillage
    use modeleturb, only: cklb1, ckle2, sigmal, xkappa
    use schemanum, only: ischema
    use codens_c
    implicit none
    integer :: l
    integer :: mnpar(ip12)
    double precision, target :: sn(ip31*ndir), vol(ip11), v(ip11,ip60), mu(ip12), mut(ip12), dist(ip12)
    double precision, target :: dvxx(ip00), dvxy(ip00), dvxz(ip00), dvyx(ip00), dvyy(ip00), dvyz(ip00)
    double precision, target :: dvzx(ip00), dvzy(ip00), dvzz(ip00), tprod(ip00)
    double precision, target :: txxf5x(ip12), txyf5y(ip12), txzf5z(ip12), tyyf6x(ip12), tyzf6y(ip12), tzzf6z(ip12)
    double precision, target :: cfke(ip13), qcxts5(ip12), qcyts6(ip12)
    double precision :: t(ip00), dtdx(ip00), dtdy(ip00), dtdz(ip00), bark(ip00)
    double precision :: tn1(ip00), tn2(ip00), tn3(ip00), tn9(ip00)
    double precision :: cmui1(ip21), cmui2(ip21), cmuj1(ip21), cmuj2(ip21), cmuk1(ip21), cmuk2(ip21)
    type(codens_block) :: b
    type(c_ptr) :: dsn, dvol, dv, dmu, dmut, ddist, dfd, dg, dtp, dcf, d5, d6, f(6), g(9)
    real(c_double) :: kc(4)
    b = codens_block_of(l)
    kc = [cklb1, ckle2, sigmal, xkappa]
    f = [c_loc(txxf5x), c_loc(txyf5y), c_loc(txzf5z), c_loc(tyyf6x), c_loc(tyzf6y), c_loc(tzzf6z)]
    g = [c_loc(dvxx), c_loc(dvxy), c_loc(dvxz), c_loc(dvyx), c_loc(dvyy), c_loc(dvyz), c_loc(dvzx), c_loc(dvzy), &
         c_loc(dvzz)]
    dsn = codens_dev_all(codens_sn_blk(sn, l), 9*nnn(l), codens_geom_mode())
    dvol = codens_dev(c_loc(vol), ip11, ip11, 1, 1, b, codens_geom_mode())
    dv = codens_dev(c_loc(v), ip11*ip60, ip11, 1, 7, b, 1)
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    ddist = codens_dev(c_loc(dist), ip12, ip12, 1, 1, b, codens_geom_mode())
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    dg = codens_bundle_in(g, 9, nnn(l), 0, nnn(l))
    dtp = codens_dev_all(c_loc(tprod), nnn(l), 1)
    dcf = codens_dev(c_loc(cfke), ip13, ip13, 1, 1, b, 1)
    d5 = codens_dev(c_loc(qcxts5), ip12, ip12, 1, 1, b, 1)
    d6 = codens_dev(c_loc(qcyts6), ip12, ip12, 1, 1, b, 1)
    call codens_check(codens_met_smklsas(b, kc, int(ischema, c_int32_t), dsn, dvol, dv, int(ip11, c_int64_t), dmu, &
         dmut, ddist, dfd, dg, dtp, dcf, d5, d6, c_null_ptr), 'met_smklsas')
    call codens_back(c_loc(cfke), ip13, 1, 1, b)
    call codens_back(c_loc(qcxts5), ip12, 1, 1, b)
    call codens_back(c_loc(qcyts6), ip12, 1, 1, b)
  end subroutine met_smklsas
end module mod_met_smklsas
