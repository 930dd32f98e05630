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
  end interface

contains

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
