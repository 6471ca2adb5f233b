! Test-infrastructure hook for the reference build (oracle/build_ref.sh).
!
! Replaces the reference's `preview` output routine (src/nhwave.F:3291-3509),
! which writes f15.6 text, with a raw little-endian FP64 dump of the full
! persistent model state, so that golden vectors keep all 53 mantissa bits.
! The driver calls it exactly where it calls `preview` (src/nhwave.F:290-296);
! run with PLOT_INTV = 0 so that it is called after every step, and list the
! steps to keep in the environment variable NHW_DUMP_STEPS (comma separated,
! e.g. "1,10,400"; unset or "all" keeps every step).
!
! File: dump_NNNNN.bin (NNNNN = RUN_STEP), stream-unformatted:
!   int32  magic(=20251202), Mloc, Nloc, Kloc, RUN_STEP, Icount
!   real64 TIME, dt
!   real64 Eta(Mloc,Nloc), D(Mloc,Nloc)
!   int32  Mask(Mloc,Nloc)
!   real64 U,V,W,DU,DV,DW (Mloc,Nloc,Kloc)
!   real64 P(Mloc,Nloc,Kloc+1), Omega(Mloc,Nloc,Kloc+1)
!   real64 Hfx0(Mloc+1,Nloc), Hfy0(Mloc,Nloc+1)
!   real64 EtaxL,EtaxR (Mloc+1,Nloc), EtayL,EtayR (Mloc,Nloc+1)
! and, with VISCOUS_FLOW, turb_NNNNN.bin (see below).
subroutine preview
  use global
  implicit none
  character(len=32) :: fname
  character(len=4096) :: steps
  character(len=16) :: tok
  integer :: iu, ln, st, p0, p1
  logical :: keep
  Icount = Icount + 1
  call get_environment_variable('NHW_DUMP_STEPS', steps, ln, st)
  keep = (st /= 0) .or. (trim(steps) == 'all')
  if (.not. keep) then
    p0 = 1
    do while (p0 <= ln)
      p1 = index(steps(p0:ln), ',')
      if (p1 == 0) then
        tok = steps(p0:ln)
        p0 = ln + 1
      else
        tok = steps(p0:p0+p1-2)
        p0 = p0 + p1
      endif
      read(tok, *) p1
      if (p1 == RUN_STEP) keep = .true.
    enddo
  endif
  if (.not. keep) return
  write(fname, '(A5,I5.5,A4)') 'dump_', RUN_STEP, '.bin'
  open(newunit=iu, file=trim(fname), access='stream', form='unformatted', status='replace')
  write(iu) 20251202, Mloc, Nloc, Kloc, RUN_STEP, Icount
  write(iu) TIME, dt
  write(iu) Eta, D
  write(iu) Mask
  write(iu) U, V, W, DU, DV, DW
  write(iu) P, Omega
  write(iu) Hfx0, Hfy0
  write(iu) EtaxL, EtaxR, EtayL, EtayR
  close(iu)
  ! viscous runs: the turbulence state in a second file, turb_NNNNN.bin:
  !   real64 Tke, Eps, DTke, DEps, CmuVt, CmuHt (Mloc,Nloc,Kloc)
  if (VISCOUS_FLOW) then
    write(fname, '(A5,I5.5,A4)') 'turb_', RUN_STEP, '.bin'
    open(newunit=iu, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(iu) Tke, Eps, DTke, DEps, CmuVt, CmuHt
    close(iu)
  endif
end subroutine preview
