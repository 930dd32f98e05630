! iso_c_binding view of include/codens.h for the drop-in Fortran routines.
!
! The records mirror the C structs; the helpers build them from the CodeNS
! module globals (maillage, proprieteflu, definition, schemanum) exactly as
! the reference routines read those globals, and move host arrays to and from
! their device mirrors.  Any non-zero status stops the run with the library's
! message, as the reference stops or aborts (zvismo.f90:122-127, chrono.f90:141).
module codens_c
  use iso_c_binding
  implicit none

  type, bind(C) :: codens_block
     integer(c_int32_t) :: npc, id1, id2, jd1, jd2, kd1, kd2, ii1, ii2, jj1, jj2, kk1, kk2, nnn, pad0, pad1
  end type codens_block

  type, bind(C) :: codens_phys
     real(c_double) :: gam, gam1, pinfl, pr, prt, cp, reynz, tnz
     integer(c_int32_t) :: iflu, pad
  end type codens_phys

  type, bind(C) :: codens_scheme
     integer(c_int32_t) :: muscl, ilim
     real(c_double) :: xk
  end type codens_scheme

  ! one wall-normal line of met_mtcorf1 (first cell 0-based)
  type, bind(C) :: codens_wall_line
     integer(c_int64_t) :: nmin
     integer(c_int32_t) :: m3min, m3max, isens3, dm3, mpn0, mpn1
  end type codens_wall_line

  ! lines of the walls of the last codens_wall_lines call, kept between calls
  ! so that their host address (the key of their device mirror) is stable
  type(codens_wall_line), allocatable, target, save :: codens_lines(:)
  integer(c_int64_t), allocatable, save :: codens_face_start(:)

  ! .true. once the caller's metrics sn and volumes vol are final: they are
  ! then uploaded whole on first use and not again.
  logical, save :: codens_geom_once = .false.

  interface
     function codens_last_error() bind(C, name="codens_last_error")
       import
       type(c_ptr) :: codens_last_error
     end function codens_last_error
     integer(c_int32_t) function codens_host_mirror(host, bytes, offset, width, pitch, rows, to_device, dev, &
          stream) bind(C, name="codens_host_mirror")
       import
       type(c_ptr), value :: host, stream
       integer(c_int64_t), value :: bytes, offset, width, pitch
       integer(c_int32_t), value :: rows, to_device
       type(c_ptr) :: dev
     end function codens_host_mirror
     integer(c_int32_t) function codens_host_fetch(host, offset, width, pitch, rows, stream) &
          bind(C, name="codens_host_fetch")
       import
       type(c_ptr), value :: host, stream
       integer(c_int64_t), value :: offset, width, pitch
       integer(c_int32_t), value :: rows
     end function codens_host_fetch
     integer(c_int32_t) function codens_host_bundle(hosts, count, bytes, offset, width, to_device, dev, stream) &
          bind(C, name="codens_host_bundle")
       import
       type(c_ptr) :: hosts(*)
       integer(c_int32_t), value :: count, to_device
       integer(c_int64_t), value :: bytes, offset, width
       type(c_ptr) :: dev
       type(c_ptr), value :: stream
     end function codens_host_bundle
     integer(c_int32_t) function codens_host_bundle_fetch(hosts, count, bytes, offset, width, stream) &
          bind(C, name="codens_host_bundle_fetch")
       import
       type(c_ptr) :: hosts(*)
       integer(c_int32_t), value :: count
       integer(c_int64_t), value :: bytes, offset, width
       type(c_ptr), value :: stream
     end function codens_host_bundle_fetch
     integer(c_int32_t) function codens_bc_coincident(nf, ghost, inner, remote, v, ldv, ps, temp, cson, stream) &
          bind(C, name="codens_bc_coincident")
       import
       integer(c_int64_t), value :: nf, ldv
       type(c_ptr), value :: ghost, inner, remote, v, ps, temp, cson, stream
     end function codens_bc_coincident
     integer(c_int32_t) function codens_bc_wall(nf, ghost, inner, v, ldv, ps, temp, cson, stream) &
          bind(C, name="codens_bc_wall")
       import
       integer(c_int64_t), value :: nf, ldv
       type(c_ptr), value :: ghost, inner, v, ps, temp, cson, stream
     end function codens_bc_wall
     integer(c_int32_t) function codens_bc_farfield(nf, ghost, inner, nrm, w0, ph, v, ldv, ps, temp, cson, stream) &
          bind(C, name="codens_bc_farfield")
       import
       integer(c_int64_t), value :: nf, ldv
       type(c_ptr), value :: ghost, inner, nrm, v, ps, temp, cson, stream
       real(c_double) :: w0(5)
       type(codens_phys) :: ph
     end function codens_bc_farfield
     integer(c_int32_t) function codens_bc_sym(nf, ghost, inner, nrm, v, ldv, stream) bind(C, name="codens_bc_sym")
       import
       integer(c_int64_t), value :: nf, ldv
       type(c_ptr), value :: ghost, inner, nrm, v, stream
     end function codens_bc_sym
     integer(c_int32_t) function codens_bc_tq_copy(nf, ghost, inner, tq, ldv, stream) &
          bind(C, name="codens_bc_tq_copy")
       import
       integer(c_int64_t), value :: nf, ldv
       type(c_ptr), value :: ghost, inner, tq, stream
     end function codens_bc_tq_copy
     integer(c_int32_t) function codens_bc_tq_coincident(nf, ghost, inner, remote, tq, ldv, stream) &
          bind(C, name="codens_bc_tq_coincident")
       import
       integer(c_int64_t), value :: nf, ldv
       type(c_ptr), value :: ghost, inner, remote, tq, stream
     end function codens_bc_tq_coincident
     integer(c_int32_t) function codens_host_residency(on, previous, stream) bind(C, name="codens_host_residency")
       import
       integer(c_int32_t), value :: on
       integer(c_int32_t) :: previous
       type(c_ptr), value :: stream
     end function codens_host_residency
     integer(c_int32_t) function codens_host_sync(host, stream) bind(C, name="codens_host_sync")
       import
       type(c_ptr), value :: host, stream
     end function codens_host_sync
     integer(c_int32_t) function codens_host_sync_all(stream) bind(C, name="codens_host_sync_all")
       import
       type(c_ptr), value :: stream
     end function codens_host_sync_all
     integer(c_int32_t) function codens_host_modified(host) bind(C, name="codens_host_modified")
       import
       type(c_ptr), value :: host
     end function codens_host_modified
     integer(c_int32_t) function codens_zvisqc(b, ph, sn, vol, v, ldv, temp, mu, mut, tq, stream) &
          bind(C, name="codens_zvisqc")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(c_ptr), value :: sn, vol, v, temp, mu, mut, tq, stream
       integer(c_int64_t), value :: ldv
     end function codens_zvisqc
     integer(c_int32_t) function codens_zpres(b, ph, v, ldv, ps, temp, cson, stream) bind(C, name="codens_zpres")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(c_ptr), value :: v, ps, temp, cson, stream
       integer(c_int64_t), value :: ldv
     end function codens_zpres
     integer(c_int32_t) function codens_zvismo(b, ph, temp, mu, nbad, stream) bind(C, name="codens_zvismo")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(c_ptr), value :: temp, mu, nbad, stream
     end function codens_zvismo
     integer(c_int32_t) function codens_chronos(b, ph, eta, ns, v, ldv, mu, mut, sn, vol, cson, dt, stream) &
          bind(C, name="codens_chronos")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       real(c_double), value :: eta
       integer(c_int32_t), value :: ns
       type(c_ptr), value :: v, mu, mut, sn, vol, cson, dt, stream
       integer(c_int64_t), value :: ldv
     end function codens_chronos
     integer(c_int32_t) function codens_sch_roe(b, ph, sc, ityprk, u, v, ldv, ff, tq, sn, ps, stream) &
          bind(C, name="codens_sch_roe")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(codens_scheme), intent(in) :: sc
       integer(c_int32_t), value :: ityprk
       type(c_ptr), value :: u, v, ff, tq, sn, ps, stream
       integer(c_int64_t), value :: ldv
     end function codens_sch_roe
     integer(c_int32_t) function codens_sch_roe_euler(b, ph, sc, ityprk, u, v, ldv, ff, sn, ps, stream) &
          bind(C, name="codens_sch_roe_euler")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(codens_scheme), intent(in) :: sc
       integer(c_int32_t), value :: ityprk
       type(c_ptr), value :: u, v, ff, sn, ps, stream
       integer(c_int64_t), value :: ldv
     end function codens_sch_roe_euler
     integer(c_int32_t) function codens_sch_expli(b, u, v, ldv, dt, vol, stream) bind(C, name="codens_sch_expli")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: u, v, dt, vol, stream
       integer(c_int64_t), value :: ldv
     end function codens_sch_expli
     integer(c_int32_t) function codens_implimf_3d(b, ph, u, dt, v, ldv, d, ff, mu, mut, lmx, sn, vol, dtpas, &
          kfmg, ityprk, ps, cson, stream) bind(C, name="codens_implimf_3d")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(c_ptr), value :: u, dt, v, d, ff, mu, mut, sn, vol, ps, cson, stream
       integer(c_int64_t), value :: ldv
       integer(c_int32_t), value :: lmx, kfmg, ityprk
       real(c_double), value :: dtpas
     end function codens_implimf_3d
     integer(c_int32_t) function codens_implimf_3d_f32(b, ph, u, dt, v, ldv, d, ff, mu, mut, lmx, sn, vol, dtpas, &
          kfmg, ityprk, ps, cson, stream) bind(C, name="codens_implimf_3d_f32")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(c_ptr), value :: u, dt, v, d, ff, mu, mut, sn, vol, ps, cson, stream
       integer(c_int64_t), value :: ldv
       integer(c_int32_t), value :: lmx, kfmg, ityprk
       real(c_double), value :: dtpas
     end function codens_implimf_3d_f32
     integer(c_int32_t) function codens_implimf_eu(b, ph, u, dt, v, ldv, nvar, d, ff, lmx, sn, vol, dt1min, kfmg, &
          ityprk, ps, cson, stream) bind(C, name="codens_implimf_eu")
       import
       type(codens_block), intent(in) :: b
       type(codens_phys), intent(in) :: ph
       type(c_ptr), value :: u, dt, v, d, ff, sn, vol, ps, cson, stream
       integer(c_int64_t), value :: ldv
       integer(c_int32_t), value :: nvar, lmx, kfmg, ityprk
       real(c_double), value :: dt1min
     end function codens_implimf_eu
     integer(c_int32_t) function codens_dissip_jameson(b, ki2, ki4, v, ldv, d, sn, ps, cson, stream) &
          bind(C, name="codens_dissip_jameson")
       import
       type(codens_block), intent(in) :: b
       real(c_double), value :: ki2, ki4
       type(c_ptr), value :: v, d, sn, ps, cson, stream
       integer(c_int64_t), value :: ldv
     end function codens_dissip_jameson
     integer(c_int32_t) function codens_met_roe2o(b, t, ldv, d, sn, stream) bind(C, name="codens_met_roe2o")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: t, d, sn, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_roe2o
     integer(c_int32_t) function codens_met_roe(b, t, ldv, d, sn, stream) bind(C, name="codens_met_roe")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: t, d, sn, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_roe
     integer(c_int32_t) function codens_teq_gradv(b, sn, vol, v, ldv, g9, stream) bind(C, name="codens_teq_gradv")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: sn, vol, v, g9, stream
       integer(c_int64_t), value :: ldv
     end function codens_teq_gradv
     integer(c_int32_t) function codens_met_prod(b, mut, v, ldv, g9, tprod, stream) bind(C, name="codens_met_prod")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: mut, v, g9, tprod, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_prod
     integer(c_int32_t) function codens_met_gradtr(b, sn, vol, v, ldv, fd, stream) bind(C, name="codens_met_gradtr")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: sn, vol, v, fd, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_gradtr
     integer(c_int32_t) function codens_met_smsa(b, tc, sn, vol, v, ldv, mu, dist, g9, fd, cfke, ts6, ts7, stream) &
          bind(C, name="codens_met_smsa")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       type(c_ptr), value :: sn, vol, v, mu, dist, g9, fd, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smsa
     integer(c_int32_t) function codens_met_smko(b, tc, v, ldv, tprod, cfke, ts6, ts7, cson, stream) &
          bind(C, name="codens_met_smko")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       type(c_ptr), value :: v, tprod, cfke, ts6, ts7, cson, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smko
     integer(c_int32_t) function codens_met_smkor(b, tc, sigmd, v, ldv, fd, g9, tprod, cfke, ts6, ts7, stream) &
          bind(C, name="codens_met_smkor")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       real(c_double), value :: sigmd
       type(c_ptr), value :: v, fd, g9, tprod, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smkor
     integer(c_int32_t) function codens_met_smch(b, cke1, cke2, utaumin, v, ldv, mu, dist, mnpar, utau, tprod, &
          ts6, ts7, stream) bind(C, name="codens_met_smch")
       import
       type(codens_block), intent(in) :: b
       real(c_double), value :: cke1, cke2, utaumin
       type(c_ptr), value :: v, mu, dist, mnpar, utau, tprod, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smch
     integer(c_int32_t) function codens_met_flud(b, kfludis, tc, v, ldv, mu, mut, fd, stream) &
          bind(C, name="codens_met_flud")
       import
       type(codens_block), intent(in) :: b
       integer(c_int32_t), value :: kfludis
       real(c_double), intent(in) :: tc(16)
       type(c_ptr), value :: v, mu, mut, fd, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_flud
     integer(c_int32_t) function codens_sch_turb(b, u, v, ldv, d, fd, ts6, ts7, sn, vol, stream) &
          bind(C, name="codens_sch_turb")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: u, v, d, fd, ts6, ts7, sn, vol, stream
       integer(c_int64_t), value :: ldv
     end function codens_sch_turb
     integer(c_int32_t) function codens_impli2_eqt_3d(b, u, ldv, dt, v, mu, mut, cfke, sn, vol, cmt, stream) &
          bind(C, name="codens_impli2_eqt_3d")
       import
       type(codens_block), intent(in) :: b
       type(c_ptr), value :: u, dt, v, mu, mut, cfke, sn, vol, stream
       integer(c_int64_t), value :: ldv
       real(c_double), value :: cmt
     end function codens_impli2_eqt_3d
     integer(c_int32_t) function codens_met_samut(b, tc, v, ldv, mu, mut, stream) bind(C, name="codens_met_samut")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       type(c_ptr), value :: v, mu, mut, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_samut
     integer(c_int32_t) function codens_met_kocmut(b, tc, kwsst, sn, vol, v, ldv, mu, mut, dist, g9, stream) &
          bind(C, name="codens_met_kocmut")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       integer(c_int32_t), value :: kwsst
       type(c_ptr), value :: sn, vol, v, mu, mut, dist, g9, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_kocmut
     integer(c_int32_t) function codens_met_smdes(b, tc, x, y, z, delta, sn, vol, v, ldv, mu, dist, g9, fd, cfke, &
          ts6, ts7, stream) bind(C, name="codens_met_smdes")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       type(c_ptr), value :: x, y, z, delta, sn, vol, v, mu, dist, g9, fd, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smdes
     integer(c_int32_t) function codens_met_smsasas(b, tc, ischema, sn, vol, v, ldv, mu, dist, g9, fd, cfke, ts6, &
          ts7, stream) bind(C, name="codens_met_smsasas")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: tc(16)
       integer(c_int32_t), value :: ischema
       type(c_ptr), value :: sn, vol, v, mu, dist, g9, fd, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smsasas
     integer(c_int32_t) function codens_met_smkl(b, kc, sn, vol, v, ldv, mu, mut, dist, fd, g9, tprod, cfke, ts6, &
          ts7, stream) bind(C, name="codens_met_smkl")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: kc(4)
       type(c_ptr), value :: sn, vol, v, mu, mut, dist, fd, g9, tprod, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smkl
     integer(c_int32_t) function codens_met_smke(b, kc, variant, lparoi, ischema, sn, vol, v, ldv, mu, mut, tprod, &
          g9, cson, bark, bare, cfke, ts6, ts7, stream) bind(C, name="codens_met_smke")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: kc(3)
       integer(c_int32_t), value :: variant, lparoi, ischema
       type(c_ptr), value :: sn, vol, v, mu, mut, tprod, g9, cson, bark, bare, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smke
     integer(c_int32_t) function codens_met_kemut(b, cmu, form, lparoi, sn, vol, v, ldv, mu, mut, dist, g9, stream) &
          bind(C, name="codens_met_kemut")
       import
       type(codens_block), intent(in) :: b
       real(c_double), value :: cmu
       integer(c_int32_t), value :: form, lparoi
       type(c_ptr), value :: sn, vol, v, mu, mut, dist, g9, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_kemut
     integer(c_int32_t) function codens_met_smklsas(b, kc, ischema, sn, vol, v, ldv, mu, mut, dist, fd, g9, tprod, &
          cfke, ts6, ts7, stream) bind(C, name="codens_met_smklsas")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: kc(4)
       integer(c_int32_t), value :: ischema
       type(c_ptr), value :: sn, vol, v, mu, mut, dist, fd, g9, tprod, cfke, ts6, ts7, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_smklsas
     integer(c_int32_t) function codens_met_klmut(b, kc, v, ldv, mu, mut, dist, stream) bind(C, name="codens_met_klmut")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: kc(4)
       type(c_ptr), value :: v, mu, mut, dist, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_klmut
     integer(c_int32_t) function codens_met_mtcorf1(nfaces, face_start, lines, imxclko, mnpar, frac, nfrac, stream) &
          bind(C, name="codens_met_mtcorf1")
       import
       integer(c_int32_t), value :: nfaces, imxclko
       integer(c_int64_t), intent(in) :: face_start(*)
       type(c_ptr), value :: lines, mnpar, frac, stream
       integer(c_int64_t), value :: nfrac
     end function codens_met_mtcorf1
     integer(c_int32_t) function codens_met_smmt(b, mc, kfracom, imxclko, v, ldv, mu, mut, dist, fd, tprod, cfke, &
          gkgo, frac, ts6, ts7, nfaces, face_start, lines, mnpar, stream) bind(C, name="codens_met_smmt")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: mc(10)
       integer(c_int32_t), value :: kfracom, imxclko, nfaces
       type(c_ptr), value :: v, mu, mut, dist, fd, tprod, cfke, gkgo, frac, ts6, ts7, lines, mnpar, stream
       integer(c_int64_t), value :: ldv
       integer(c_int64_t), intent(in) :: face_start(*)
     end function codens_met_smmt
     integer(c_int32_t) function codens_met_smmtr(b, mc, imxclko, v, ldv, mu, mut, dist, fd, g9, tprod, cfke, &
          frac, ts6, ts7, nfaces, face_start, lines, mnpar, stream) bind(C, name="codens_met_smmtr")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: mc(10)
       integer(c_int32_t), value :: imxclko, nfaces
       type(c_ptr), value :: v, mu, mut, dist, fd, g9, tprod, cfke, frac, ts6, ts7, lines, mnpar, stream
       integer(c_int64_t), value :: ldv
       integer(c_int64_t), intent(in) :: face_start(*)
     end function codens_met_smmtr
     integer(c_int32_t) function codens_met_fludmt(b, mc, kfracom, mu, mut, frac, fd, ldv, stream) &
          bind(C, name="codens_met_fludmt")
       import
       type(codens_block), intent(in) :: b
       real(c_double), intent(in) :: mc(10)
       integer(c_int32_t), value :: kfracom
       type(c_ptr), value :: mu, mut, frac, fd, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_fludmt
     integer(c_int32_t) function codens_met_komut(b, betae, komsst, sn, vol, v, ldv, mu, mut, dist, g9, stream) &
          bind(C, name="codens_met_komut")
       import
       type(codens_block), intent(in) :: b
       real(c_double), value :: betae
       integer(c_int32_t), value :: komsst
       type(c_ptr), value :: sn, vol, v, mu, mut, dist, g9, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_komut
     integer(c_int32_t) function codens_met_cut(b, which, epsk, epse, v, ldv, stream) bind(C, name="codens_met_cut")
       import
       type(codens_block), intent(in) :: b
       integer(c_int32_t), value :: which
       real(c_double), value :: epsk, epse
       type(c_ptr), value :: v, stream
       integer(c_int64_t), value :: ldv
     end function codens_met_cut
  end interface

contains

  ! device address p advanced by `bytes`
  function codens_ptr_off(p, bytes) result(q)
    type(c_ptr), intent(in) :: p
    integer(c_int64_t), intent(in) :: bytes
    type(c_ptr) :: q
    integer(c_intptr_t) :: a
    a = transfer(p, a)
    a = a + bytes
    q = transfer(a, q)
  end function codens_ptr_off

  integer function codens_geom_mode()
    codens_geom_mode = 1
    if (codens_geom_once) codens_geom_mode = 2
  end function codens_geom_mode

  ! The kernels implement the 3-D forms (equat(3:4) = '3d') only.
  subroutine codens_need_3d(equat, what)
    character(len=*), intent(in) :: equat, what
    if (equat(3:4) .ne. '3d') then
       write(*, '(a,": only the 3-D form runs on the GPU, equat = ",a)') what, equat
       error stop 'codens: unsupported equation set'
    endif
  end subroutine codens_need_3d

  ! extents of block l from module maillage (crdms.f90:126-153)
  function codens_block_of(l) result(b)
    use maillage
    integer, intent(in) :: l
    type(codens_block) :: b
    b%npc = npc(l)
    b%id1 = id1(l); b%id2 = id2(l); b%jd1 = jd1(l); b%jd2 = jd2(l); b%kd1 = kd1(l); b%kd2 = kd2(l)
    b%ii1 = ii1(l); b%ii2 = ii2(l); b%jj1 = jj1(l); b%jj2 = jj2(l); b%kk1 = kk1(l); b%kk2 = kk2(l)
    b%nnn = nnn(l)
    b%pad0 = 0; b%pad1 = 0
  end function codens_block_of

  ! gas constants from modules proprieteflu and definition
  function codens_phys_now() result(ph)
    use proprieteflu
    use definition
    type(codens_phys) :: ph
    ph%gam = gam; ph%gam1 = gam1; ph%pinfl = pinfl; ph%pr = pr; ph%prt = prt; ph%cp = cp
    ph%reynz = reynz; ph%tnz = tnz; ph%iflu = iflu; ph%pad = 0
  end function codens_phys_now

  ! MUSCL switches from module schemanum
  function codens_scheme_now() result(sc)
    use schemanum
    type(codens_scheme) :: sc
    sc%muscl = muscl; sc%ilim = ilim; sc%xk = xk
  end function codens_scheme_now

  ! turbulence constants of module modeleturb in the order of the C-ABI
  function codens_tc_now() result(tc)
    use modeleturb
    real(c_double) :: tc(16)
    tc = [cb1, cb2, sigma, kappa, cw1, cw2, cw3, cv1, ct3, ct4, betas, beta, sigmk, sigmw, alfak, alfae]
  end function codens_tc_now

  ! offset of block b's metric slice in the global sn(ip31*ndir) that some
  ! routines receive whole (they find npsn = ndir*npfb(l)+1 themselves)
  function codens_sn_blk(sn, l) result(p)
    use maillage
    use para_fige
    double precision, target, intent(in) :: sn(*)
    integer, intent(in) :: l
    type(c_ptr) :: p
    p = c_loc(sn(ndir*npfb(l) + 1))
  end function codens_sn_blk

  ! n arrays of length len (each) as one device bundle with slot stride
  ! len; the window [lo, hi) of element indices (0-based) of each is copied
  function codens_bundle_in(arrs, cnt, len, lo, hi) result(dev)
    type(c_ptr), intent(in) :: arrs(*)
    integer, intent(in) :: cnt, len, lo, hi
    type(c_ptr) :: dev
    call codens_check(codens_host_bundle(arrs, int(cnt, c_int32_t), 8_c_int64_t*len, 8_c_int64_t*lo, &
         8_c_int64_t*(hi - lo), 1_c_int32_t, dev, c_null_ptr), 'codens_bundle_in')
  end function codens_bundle_in

  subroutine codens_bundle_back(arrs, cnt, len, lo, hi)
    type(c_ptr), intent(in) :: arrs(*)
    integer, intent(in) :: cnt, len, lo, hi
    call codens_check(codens_host_bundle_fetch(arrs, int(cnt, c_int32_t), 8_c_int64_t*len, 8_c_int64_t*lo, &
         8_c_int64_t*(hi - lo), c_null_ptr), 'codens_bundle_back')
  end subroutine codens_bundle_back

  ! The met_flud* family on block l: kfludis selects the diffusivity
  ! (met_num.f90:528-556); the six arrays go to the device as one bundle of
  ! the block's window and come back transformed.  s is read for kfludis = 2
  ! only, mut for 1 and 5 only.
  subroutine codens_flud(kfludis, l, s, mu, mut, fd5x, fd5y, fd5z, fd6x, fd6y, fd6z, what)
    use para_var
    use maillage
    integer, intent(in) :: kfludis, l
    double precision, target :: s(*), mu(ip12), mut(ip12)
    double precision, target :: fd5x(ip12), fd5y(ip12), fd5z(ip12), fd6x(ip12), fd6y(ip12), fd6z(ip12)
    character(len=*), intent(in) :: what
    type(codens_block) :: b
    type(c_ptr) :: dv, dmu, dmut, dfd, f(6)
    b = codens_block_of(l)
    f = [c_loc(fd5x), c_loc(fd5y), c_loc(fd5z), c_loc(fd6x), c_loc(fd6y), c_loc(fd6z)]
    dv = c_null_ptr
    dmut = c_null_ptr
    if (kfludis .eq. 2) then
       dv = codens_dev(c_loc(s), ip11*ip60, ip11, 1, 7, b, 1)
    else
       dmut = codens_dev(c_loc(mut), ip12, ip12, 1, 1, b, 1)
    endif
    dmu = codens_dev(c_loc(mu), ip12, ip12, 1, 1, b, 1)
    dfd = codens_bundle_in(f, 6, ip12, npc(l), npc(l) + nnn(l))
    call codens_check(codens_met_flud(b, int(kfludis, c_int32_t), codens_tc_now(), dv, int(ip11, c_int64_t), dmu, &
         dmut, dfd, c_null_ptr), what)
    call codens_bundle_back(f, 6, ip12, npc(l), npc(l) + nnn(l))
  end subroutine codens_flud

  ! at_cutke limiters on block l: which = 1 met_cut, 2 met_cutsa
  subroutine codens_cut(which, l, s, epsk, epse)
    use para_var
    integer, intent(in) :: which, l
    double precision, target :: s(ip11,ip60)
    double precision, intent(in) :: epsk, epse
    type(codens_block) :: b
    type(c_ptr) :: dv
    b = codens_block_of(l)
    dv = codens_dev(c_loc(s), ip11*ip60, ip11, 6, 2, b, 1)
    call codens_check(codens_met_cut(b, int(which, c_int32_t), epsk, epse, dv, int(ip11, c_int64_t), c_null_ptr), &
         'met_cut')
    call codens_back(c_loc(s), ip11, 6, 2, b)
  end subroutine codens_cut

  ! Menter constants in the C-ABI's order (module modeleturb)
  function codens_menter_now() result(mc)
    use modeleturb
    real(c_double) :: mc(10)
    mc = [betae, okappa, sigme1, sigma1, beta1, wsig1, sigme2, sigma2, beta2, wsig2]
  end function codens_menter_now

  ! The wall-normal lines met_mtcorf1 walks (met_mtcorf1.f90:74-88): every
  ! point of every wall of the list lbdko on block l (all blocks for l = 0),
  ! from atindnor; into codens_lines / codens_face_start, nf walls.
  subroutine codens_wall_lines(l, ncin, nf)
    use para_var
    use boundary
    use mod_atindnor
    integer, intent(in) :: l, ncin(*)
    integer, intent(out) :: nf
    integer :: mpar, mfbi, m, nl, m10, m20, m30, m1min, m1max, m2min, m2max, m3min, m3max, dm1, dm2, dm3, isens3
    nl = 0
    nf = 0
    do mpar = 1, nbdko
       mfbi = lbdko(mpar)
       if (l .eq. 0 .or. l .eq. ndlb(mfbi)) then
          nl = nl + mmb(mfbi)
          nf = nf + 1
       endif
    enddo
    if (.not. allocated(codens_lines)) allocate(codens_lines(max(nl, 1)), codens_face_start(nf + 1))
    if (size(codens_lines) .lt. nl .or. size(codens_face_start) .lt. nf + 1) then
       deallocate(codens_lines, codens_face_start)
       allocate(codens_lines(max(nl, 1)), codens_face_start(nf + 1))
    endif
    nl = 0
    nf = 0
    codens_face_start(1) = 0
    do mpar = 1, nbdko
       mfbi = lbdko(mpar)
       if (l .eq. 0 .or. l .eq. ndlb(mfbi)) then
          call atindnor(mfbi, m10, m20, m30, m1min, m1max, m2min, m2max, m3min, m3max, dm1, dm2, dm3, isens3)
          do m = 1, mmb(mfbi)
             nl = nl + 1
             codens_lines(nl) = codens_wall_line(int(ncin(mpb(mfbi) + m) - 1, c_int64_t), m3min, m3max, isens3, &
                  dm3, mpn(mfbi), mpn(mfbi) + mmb(mfbi))
          enddo
          nf = nf + 1
          codens_face_start(nf + 1) = nl
       endif
    enddo
  end subroutine codens_wall_lines

  ! device copy of the lines of codens_wall_lines
  function codens_lines_dev() result(dev)
    type(c_ptr) :: dev
    integer(c_int64_t) :: nb
    nb = 32_c_int64_t * size(codens_lines)
    call codens_check(codens_host_mirror(c_loc(codens_lines), nb, 0_c_int64_t, nb, nb, 1_c_int32_t, 1_c_int32_t, &
         dev, c_null_ptr), 'codens_lines_dev')
  end function codens_lines_dev

  subroutine codens_check(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    character(len=512) :: text
    if (rc == 0) return
    call c_f_pointer(codens_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    enddo
    write(*, '(a,": ",a)') what, trim(text)
    error stop 'codens: GPU call failed'
  end subroutine codens_check

  ! Device mirror of host array `a` of `n` doubles.  Window: rows
  ! row0..row0+rows-1 of a(ld,*), each restricted to block b's slice
  ! [npc, npc+nnn).  mode: 0 no copy (pure output), 1 copy the window in,
  ! 2 constant array (copied whole on first use only).
  function codens_dev(a, n, ld, row0, rows, b, mode) result(dev)
    type(c_ptr), intent(in) :: a
    integer, intent(in) :: n, ld, row0, rows, mode
    type(codens_block), intent(in) :: b
    type(c_ptr) :: dev
    call codens_check(codens_host_mirror(a, 8_c_int64_t*n, 8_c_int64_t*(b%npc + int(row0-1,c_int64_t)*ld), &
         8_c_int64_t*b%nnn, 8_c_int64_t*ld, int(rows,c_int32_t), int(mode,c_int32_t), dev, c_null_ptr), &
         'codens_dev')
  end function codens_dev

  ! Whole-array mirror of `a` (n doubles), e.g. one block's metric slice.
  function codens_dev_all(a, n, mode) result(dev)
    type(c_ptr), intent(in) :: a
    integer, intent(in) :: n, mode
    type(c_ptr) :: dev
    call codens_check(codens_host_mirror(a, 8_c_int64_t*n, 0_c_int64_t, 8_c_int64_t*n, 8_c_int64_t*n, &
         1_c_int32_t, int(mode,c_int32_t), dev, c_null_ptr), 'codens_dev_all')
  end function codens_dev_all

  ! Window of codens_dev back into the host array (synchronous).
  subroutine codens_back(a, ld, row0, rows, b)
    type(c_ptr), intent(in) :: a
    integer, intent(in) :: ld, row0, rows
    type(codens_block), intent(in) :: b
    call codens_check(codens_host_fetch(a, 8_c_int64_t*(b%npc + int(row0-1,c_int64_t)*ld), 8_c_int64_t*b%nnn, &
         8_c_int64_t*ld, int(rows,c_int32_t), c_null_ptr), 'codens_back')
  end subroutine codens_back

  ! Residency hooks (codens_host_residency): host code about to read array
  ! a, and host code that has written a.  With residency off they cost a
  ! lookup.
  subroutine codens_sync(a)
    type(c_ptr), intent(in) :: a
    call codens_check(codens_host_sync(a, c_null_ptr), 'codens_sync')
  end subroutine codens_sync

  subroutine codens_modified(a)
    type(c_ptr), intent(in) :: a
    call codens_check(codens_host_modified(a), 'codens_modified')
  end subroutine codens_modified

  ! The 9 viscous arrays toxx..qcz (each n doubles) as one device bundle
  ! with slot stride n, block b's slice copied in.
  function codens_tq_in(tq, n, b) result(dev)
    type(c_ptr), intent(in) :: tq(9)
    integer, intent(in) :: n
    type(codens_block), intent(in) :: b
    type(c_ptr) :: dev
    call codens_check(codens_host_bundle(tq, 9_c_int32_t, 8_c_int64_t*n, 8_c_int64_t*b%npc, 8_c_int64_t*b%nnn, &
         1_c_int32_t, dev, c_null_ptr), 'codens_tq_in')
  end function codens_tq_in

  subroutine codens_tq_back(tq, n, b)
    type(c_ptr), intent(in) :: tq(9)
    integer, intent(in) :: n
    type(codens_block), intent(in) :: b
    call codens_check(codens_host_bundle_fetch(tq, 9_c_int32_t, 8_c_int64_t*n, 8_c_int64_t*b%npc, &
         8_c_int64_t*b%nnn, c_null_ptr), 'codens_tq_back')
  end subroutine codens_tq_back

end module codens_c
