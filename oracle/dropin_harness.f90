! Checker entry points that exist only in the drop-in build
! (oracle/_ref/libcodens_dropin.so): they call routines whose reference
! originals are not in the checker build, so only their drop-in form can be
! driven.  TEST INFRASTRUCTURE ONLY, like ref_harness.f90.
module codens_dropin_harness
  use iso_c_binding
  use para_var
  use para_fige
  implicit none
contains
  ! rbord.f90:4 over the frontiers of ref_set_boundaries: kinds(f) = 0
  ! coincident ('rc'), 1 wall ('para'), 2 far field ('nrd'), 3 symmetry
  ! ('sym'); blocks(f) the frontier's block (1-based); the coincident list
  ! nfbc = frc(1:nc); every far field gets the state of Mach number mach and
  ! incidence alpha, beta (degrees) against the stagnation state roa, aa, pa
  ! (utnrd.f90:41-53); nxn, nyn, nzn: the inward unit normals of every face.
  subroutine ref_rbord(nfr, kinds, blocks, nc, frc, ncbd, ncin, nxn, nyn, nzn, v, ps, temp, cson, mach, alpha, &
       beta, roa, aa, pa, y, z) bind(C, name="ref_rbord")
    use boundary
    use definition
    use schemanum
    use mod_rbord
    integer(c_int), value :: nfr, nc
    integer(c_int) :: kinds(*), blocks(*), frc(*), ncbd(*), ncin(*)
    real(c_double), value :: mach, alpha, beta, roa, aa, pa
    real(c_double) :: nxn(*), nyn(*), nzn(*), v(*), ps(*), temp(*), cson(*), y(*), z(*)
    ! the routine's scratch and unused inputs, kept between calls (as a
    ! driver keeps its arrays) and remade when the layout grows
    real(c_double), allocatable, save :: bceqt(:, :), sc(:, :), xyz(:)
    real(c_double) :: sn(1), vol(1), dt(1)
    integer :: mnc(1), mnr(1), f
    real(c_double) :: xnr(1), ynr(1), znr(1)
    character(len=4), parameter :: names(0:3) = ['rc  ', 'para', 'nrd ', 'sym ']
    if (allocated(cl)) deallocate(cl)
    if (allocated(nfbc)) deallocate(nfbc)
    allocate(cl(nfr), nfbc(max(nc, 1)))
    if (allocated(bceqt)) then
       if (size(bceqt, 1) < ip41 .or. size(sc, 1) < ip40 .or. size(xyz) < ip21) deallocate(bceqt, sc, xyz)
    endif
    if (.not. allocated(bceqt)) then
       allocate(bceqt(ip41, neqt), sc(ip40, 18), xyz(ip21))
       bceqt = 0.; sc = 0.; xyz = 0.
    endif
    sn = 0.; vol = 0.; dt = 0.
    do f = 1, nfr
       cl(f) = names(kinds(f))
       nba(f) = f
       ndlb(f) = blocks(f)
       if (kinds(f) == 2) then
          bceqt(mpb(f) + 1, 1) = mach
          bceqt(mpb(f) + 1, 2) = alpha
          bceqt(mpb(f) + 1, 3) = beta
       endif
    enddo
    do f = 1, nc
       nfbc(f) = frc(f)
    enddo
    mtbx = nfr; mtcx = nc; mtb = 0; kprec = 0
    roa1 = roa; aa1 = aa; pa1 = pa; omg = 0.
    mnc = 0; mnr = 0; xnr = 0.; ynr = 0.; znr = 0.
    call rbord(0, 1, mnc, ncin, mnr, xnr, ynr, znr, nxn, nyn, nzn, ncbd, sn, vol, v, dt, bceqt, &
         sc(1,1), sc(1,2), sc(1,3), sc(1,4), sc(1,5), xyz, y, z, sc(1,6), sc(1,7), sc(1,8), sc(1,9), sc(1,10), &
         sc(1,11), sc(1,12), 1, sc(1,13), sc(1,14), sc(1,15), sc(1,16), sc(1,17), sc(1,18), sc(1,1), sc(1,2), &
         sc(1,3), sc(1,4), sc(1,5), sc(1,6), sc(1,7), ps, temp, cson)
  end subroutine ref_rbord
end module codens_dropin_harness
