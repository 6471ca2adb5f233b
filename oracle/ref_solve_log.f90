! Test-infrastructure hook for the reference build (oracle/build_ref.sh).
!
! Wraps NSPCG's driver (src/nspcg.f), which the reference's poisson_solver
! calls once per Poisson solve (src/pressure.F:122-130), to log what each
! solve achieves as a true residual: before and after the call it forms
! ||Rhs - A Phi|| with A the diagonal-storage matrix that generate_coef_rhs
! assembled (Coef(1:n,1:maxnz), offsets JCoef(1:maxnz), src/pressure.F:817-831)
! from copies taken before the call.  The original driver is linked under
! the name nspcg_orig (a renamed copy of its object file); the solve itself
! is unchanged.
!
! File nhw_solves.txt in the run directory, one line per solve, appended:
!   RUN_STEP  n  ||Rhs||  ||Rhs - A Phi0||  ||Rhs - A Phi||  ier
! (Phi0: the initial guess the reference passes, the previous P).
subroutine nspcg(precon, accel, ndim, mdim, n, maxnz, coef, jcoef, p, ip, u, ubar, rhs, wksp, iwksp, &
                 nw, inw, iparm, rparm, ier)
  use global, only: SP, RUN_STEP
  implicit none
  external :: precon, accel
  integer :: ndim, mdim, n, maxnz, nw, inw, ier
  integer :: jcoef(*), p(*), ip(*), iwksp(*), iparm(30)
  real(SP) :: coef(ndim, *), u(*), ubar(*), rhs(*), wksp(*), rparm(30)
  real(8), allocatable :: c0(:, :), b0(:)
  integer, allocatable :: j0(:)
  real(8) :: r0, r1, bn
  integer :: iu
  allocate(c0(n, maxnz), b0(n), j0(maxnz))
  c0 = coef(1:n, 1:maxnz)
  j0 = jcoef(1:maxnz)
  b0 = rhs(1:n)
  bn = sqrt(sum(b0 * b0))
  r0 = resid_norm(u)
  call nspcg_orig(precon, accel, ndim, mdim, n, maxnz, coef, jcoef, p, ip, u, ubar, rhs, wksp, iwksp, &
                  nw, inw, iparm, rparm, ier)
  r1 = resid_norm(u)
  open(newunit=iu, file='nhw_solves.txt', position='append', action='write')
  write(iu, '(i9,1x,i10,3(1x,es24.16e3),1x,i6)') RUN_STEP, n, bn, r0, r1, ier
  close(iu)
  deallocate(c0, b0, j0)
contains
  real(8) function resid_norm(x)
    real(SP), intent(in) :: x(*)
    integer :: i, d, q
    real(8) :: acc, s
    s = 0.0d0
    do i = 1, n
      acc = b0(i)
      do d = 1, maxnz
        q = i + j0(d)
        if (q >= 1 .and. q <= n) acc = acc - c0(i, d) * x(q)
      end do
      s = s + acc * acc
    end do
    resid_norm = sqrt(s)
  end function resid_norm
end subroutine nspcg
