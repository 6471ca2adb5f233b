! ISO_C_BINDING drop-in for the NHWAVE time-stepping hot path.
!
! Defines the reference's external subroutines
!   estimate_dt, update_mask, update_wind, update_vars, source_terms,
!   fluxes, eval_duvw(ISTEP), poisson_solver, update_wave_bc, sponge_damping,
!   eval_turb(ISTEP)
! (src/nhwave.F:3694, 5403, 5333, 5459, 1686, 539, 5514, 6462, 6544;
! src/source_terms.F:1; src/pressure.F:63) with the same names and
! arguments, so that linking this
! object ahead of weakened reference objects routes the driver's calls
! (src/nhwave.F:127-186) to libnhwave_amd.  No reference source is edited.
!
! State ownership: at the first call the GLOBAL arrays are uploaded to a
! device context and stay resident; the host copies go stale until a
! download.  The driver reads host fields only in statistics, probes,
! preview and max_min_property (src/nhwave.F:268-296), so after the last RK
! stage of a step the shim downloads exactly when one of those will fire,
! evaluating the driver's own counters with the same expressions.  TIME, dt
! and RUN_STEP are mirrored to the host after every estimate_dt, which also
! writes the reference's log line (src/nhwave.F:3852-3856).
!
! Errors: any non-zero return from the library writes the message to the log
! unit 3 and stops, as the reference stops on its fatal paths
! (src/nhwave.F:3840-3850).
!
! Environment:
!   NHW_SOLVER=line        use the line-preconditioned solver instead of
!                          multigrid
!   NHW_SNAPSHOT_STEPS=a,b full-precision binary state snapshots after the
!                          listed steps (file nhw_snap_NNNNN.bin, layout below)
!
! Snapshot layout (stream, little endian):
!   int32  20251202, Mloc, Nloc, Kloc, RUN_STEP, Icount
!   real64 TIME, dt
!   real64 Eta(Mloc,Nloc), D(Mloc,Nloc); int32 Mask(Mloc,Nloc)
!   real64 U,V,W,DU,DV,DW (Mloc,Nloc,Kloc); P, Omega (Mloc,Nloc,Kloc+1)
!   real64 Hfx0(Mloc+1,Nloc), Hfy0(Mloc,Nloc+1)
!   real64 EtaxL,EtaxR (Mloc+1,Nloc), EtayL,EtayR (Mloc,Nloc+1)

module nhw_shim
  use iso_c_binding
  implicit none
  private
  public :: nhw_shim_ctx, nhw_shim_start, nhw_shim_check, nhw_shim_end_of_step, nhw_shim_mirror_time

  type, bind(C) :: nhw_config_t
    integer(c_int32_t) :: mglob, nglob, kglob
    real(c_double) :: dx, dy, cfl, min_dep, dt_ini, dt_min, dt_max, froude_cap
    integer(c_int32_t) :: froude_cap_on
    real(c_double) :: visc, fcor, wind_u, wind_v
    integer(c_int32_t) :: time_order, non_hydro, hllc
    integer(c_int32_t) :: bc_x0, bc_xn, bc_y0, bc_yn, bc_z0
    integer(c_int32_t) :: itmax
    real(c_double) :: tol
    integer(c_int32_t) :: precision, solver
    integer(c_int32_t) :: wavemaker
    real(c_double) :: amp_wave, per_wave, dep_wave, theta_wave, cur_wave, tramp
    integer(c_int32_t) :: sponge_on
    real(c_double) :: sponge_w, sponge_e, sponge_s, sponge_n
    integer(c_int32_t) :: viscous, rng
    real(c_double) :: cvs, chs, schmidt, visc_number
    integer(c_int32_t) :: ibot
    real(c_double) :: cd0, zob
    integer(c_int32_t) :: turb
  end type

  ! same member order as nhw_fields in include/nhwave_amd.h
  type, bind(C) :: nhw_fields_t
    type(c_ptr) :: Eta = c_null_ptr, Eta0 = c_null_ptr, D = c_null_ptr, D0 = c_null_ptr, &
                   Hc = c_null_ptr, DelxH = c_null_ptr, DelyH = c_null_ptr, &
                   DelxEta = c_null_ptr, DelyEta = c_null_ptr
    type(c_ptr) :: Mask = c_null_ptr, Mask_Struct = c_null_ptr, Mask9 = c_null_ptr
    type(c_ptr) :: Hfx = c_null_ptr, Hfx0 = c_null_ptr, EtaxL = c_null_ptr, EtaxR = c_null_ptr
    type(c_ptr) :: Hfy = c_null_ptr, Hfy0 = c_null_ptr, EtayL = c_null_ptr, EtayR = c_null_ptr
    type(c_ptr) :: U = c_null_ptr, V = c_null_ptr, W = c_null_ptr, DU = c_null_ptr, &
                   DV = c_null_ptr, DW = c_null_ptr, DU0 = c_null_ptr, DV0 = c_null_ptr, &
                   DW0 = c_null_ptr
    type(c_ptr) :: P = c_null_ptr, Omega = c_null_ptr
    type(c_ptr) :: SourceX = c_null_ptr, SourceY = c_null_ptr, Wsx = c_null_ptr, Wsy = c_null_ptr
    type(c_ptr) :: Ex = c_null_ptr, Ey = c_null_ptr, Fx = c_null_ptr, Fy = c_null_ptr, &
                   Gx = c_null_ptr, Gy = c_null_ptr, Hx = c_null_ptr, Hy = c_null_ptr
    type(c_ptr) :: Fz = c_null_ptr, Gz = c_null_ptr, Hz = c_null_ptr
    type(c_ptr) :: CmuVt = c_null_ptr, Tke = c_null_ptr, Eps = c_null_ptr, DTke = c_null_ptr, &
                   DEps = c_null_ptr
  end type

  interface
    function c_last_error() bind(C, name='nhw_last_error')
      import :: c_ptr
      type(c_ptr) :: c_last_error
    end function
    integer(c_int) function c_create(cfg, ctx) bind(C, name='nhw_create')
      import :: c_int, c_ptr, nhw_config_t
      type(nhw_config_t), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function c_set_grid(ctx, dsig, sig, sigc) bind(C, name='nhw_set_grid')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dsig, sig, sigc
    end function
    integer(c_int) function c_upload(ctx, f) bind(C, name='nhw_upload')
      import :: c_int, c_ptr, nhw_fields_t
      type(c_ptr), value :: ctx
      type(nhw_fields_t), intent(in) :: f
    end function
    integer(c_int) function c_download(ctx, f) bind(C, name='nhw_download')
      import :: c_int, c_ptr, nhw_fields_t
      type(c_ptr), value :: ctx
      type(nhw_fields_t), intent(in) :: f
    end function
    integer(c_int) function c_set_time(ctx, time, dt, run_step) bind(C, name='nhw_set_time')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: time, dt
      integer(c_int32_t), value :: run_step
    end function
    integer(c_int) function c_get_time(ctx, time, dt, run_step) bind(C, name='nhw_get_time')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time, dt
      integer(c_int32_t), intent(out) :: run_step
    end function
    integer(c_int) function c_set_internal(ctx, xw, xe, ys, yn, nf, nd, freq, dire, spc, phs) &
        bind(C, name='nhw_set_internal_wavemaker')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: xw, xe, ys, yn
      integer(c_int32_t), value :: nf, nd
      real(c_double), intent(in) :: freq(*), dire(*), spc(*), phs(*)
    end function
  end interface

  type(c_ptr), save :: ctx = c_null_ptr

contains

  function nhw_shim_ctx() result(c)
    type(c_ptr) :: c
    call nhw_shim_start()
    c = ctx
  end function

  subroutine nhw_shim_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    if (rc == 0) return
    text = ''
    call c_f_pointer(c_last_error(), msg, [512])
    do n = 1, 512
      if (msg(n) == c_null_char) exit
      text(n:n) = msg(n)
    enddo
    write(3,*) 'libnhwave_amd: ', what, ' failed (', rc, '): ', trim(text)
    write(*,*) 'libnhwave_amd: ', what, ' failed (', rc, '): ', trim(text)
    stop
  end subroutine

  ! address of a contiguous GLOBAL array (allocatable module arrays carry no
  ! TARGET attribute; a contiguous actual is passed without a copy)
  function ad(a) result(p)
    real(c_double), target, intent(in) :: a(*)
    type(c_ptr) :: p
    p = c_loc(a(1))
  end function
  function ai(a) result(p)
    integer(c_int32_t), target, intent(in) :: a(*)
    type(c_ptr) :: p
    p = c_loc(a(1))
  end function

  ! every GLOBAL array the library owns, as host pointers
  subroutine all_fields(f)
    use global
    type(nhw_fields_t), intent(out) :: f
    f%Eta = ad(Eta); f%Eta0 = ad(Eta0); f%D = ad(D); f%D0 = ad(D0)
    f%Hc = ad(Hc); f%DelxH = ad(DelxH); f%DelyH = ad(DelyH)
    f%DelxEta = ad(DelxEta); f%DelyEta = ad(DelyEta)
    f%Mask = ai(Mask); f%Mask_Struct = ai(Mask_Struct); f%Mask9 = ai(Mask9)
    f%Hfx = ad(Hfx); f%Hfx0 = ad(Hfx0); f%EtaxL = ad(EtaxL); f%EtaxR = ad(EtaxR)
    f%Hfy = ad(Hfy); f%Hfy0 = ad(Hfy0); f%EtayL = ad(EtayL); f%EtayR = ad(EtayR)
    f%U = ad(U); f%V = ad(V); f%W = ad(W)
    f%DU = ad(DU); f%DV = ad(DV); f%DW = ad(DW)
    f%DU0 = ad(DU0); f%DV0 = ad(DV0); f%DW0 = ad(DW0)
    f%P = ad(P); f%Omega = ad(Omega)
    f%SourceX = ad(SourceX); f%SourceY = ad(SourceY); f%Wsx = ad(Wsx); f%Wsy = ad(Wsy)
    f%Ex = ad(Ex); f%Ey = ad(Ey); f%Fx = ad(Fx); f%Fy = ad(Fy)
    f%Gx = ad(Gx); f%Gy = ad(Gy); f%Hx = ad(Hx); f%Hy = ad(Hy)
    f%Fz = ad(Fz); f%Gz = ad(Gz); f%Hz = ad(Hz)
    if (VISCOUS_FLOW .and. IVturb == 10) then  ! the k-epsilon state
      f%CmuVt = ad(CmuVt); f%Tke = ad(Tke); f%Eps = ad(Eps); f%DTke = ad(DTke); f%DEps = ad(DEps)
    endif
  end subroutine

  ! creates the device context from GLOBAL at the first substituted call
  subroutine nhw_shim_start()
    use global
    type(nhw_config_t) :: cfg
    type(nhw_fields_t) :: f
    character(len=16) :: sv
    integer :: ln, st
    if (c_associated(ctx)) return
    if (WAVE_AVERAGE_ON .or. (VISCOUS_FLOW .and. .not. ((IVturb == 1 .and. IHturb == 1) .or. &
                                                       (IVturb == 10 .and. IHturb >= 10)))) then
      write(3,*) 'libnhwave_amd: WAVE_AVERAGE_ON and VISCOUS_FLOW with turbulence models other than', &
                 ' IVTURB = IHTURB = 1 or IVTURB = 10 with IHTURB >= 10 are not on the GPU path'
      write(*,*) 'libnhwave_amd: WAVE_AVERAGE_ON and VISCOUS_FLOW with turbulence models other than', &
                 ' IVTURB = IHTURB = 1 or IVTURB = 10 with IHTURB >= 10 are not on the GPU path'
      stop
    endif
    cfg%viscous = merge(1, 0, VISCOUS_FLOW)
    cfg%turb = merge(10, 1, VISCOUS_FLOW .and. IVturb == 10)
    cfg%rng = merge(1, 0, RNG)
    cfg%cvs = Cvs; cfg%chs = Chs; cfg%schmidt = Schmidt; cfg%visc_number = VISCOUS_NUMBER
    cfg%ibot = Ibot; cfg%cd0 = Cd0; cfg%zob = Zob
    select case (WaveMaker(1:7))
    case ('LEF_SOL')
      cfg%wavemaker = 1
    case ('LEF_LIN')
      cfg%wavemaker = 2
    case ('LEF_STK')
      cfg%wavemaker = 3
    case ('LEF_CON')
      cfg%wavemaker = 4
    case ('INT_IRR')
      cfg%wavemaker = 5
    case ('FLUX_LR')
      cfg%wavemaker = 6
    case default
      cfg%wavemaker = 0
      if (WaveMaker(1:4) /= 'NONE') then
        write(3,*) 'libnhwave_amd: WAVEMAKER ', trim(WaveMaker), ' is not on the GPU path'
        write(*,*) 'libnhwave_amd: WAVEMAKER ', trim(WaveMaker), ' is not on the GPU path'
        stop
      endif
    end select
    cfg%amp_wave = 0.0_c_double; cfg%per_wave = 1.0_c_double; cfg%dep_wave = 1.0_c_double
    cfg%theta_wave = 0.0_c_double; cfg%cur_wave = 0.0_c_double
    if (cfg%wavemaker /= 0) then
      cfg%amp_wave = Amp_Wave; cfg%per_wave = Per_Wave; cfg%dep_wave = Dep_Wave
      cfg%theta_wave = Theta_Wave; cfg%cur_wave = Cur_Wave
    endif
    cfg%tramp = TRamp
    cfg%sponge_on = merge(1, 0, SPONGE_ON)
    cfg%sponge_w = 0.0_c_double; cfg%sponge_e = 0.0_c_double
    cfg%sponge_s = 0.0_c_double; cfg%sponge_n = 0.0_c_double
    if (SPONGE_ON) then
      cfg%sponge_w = Sponge_West_Width; cfg%sponge_e = Sponge_East_Width
      cfg%sponge_s = Sponge_South_Width; cfg%sponge_n = Sponge_North_Width
    endif
    cfg%mglob = Mglob; cfg%nglob = Nglob; cfg%kglob = Kglob
    cfg%dx = dx; cfg%dy = dy; cfg%cfl = CFL; cfg%min_dep = MinDep
    cfg%dt_ini = dt_ini; cfg%dt_min = dt_min; cfg%dt_max = dt_max
#if defined (FROUDE_CAP)
    cfg%froude_cap = FROUDECAP; cfg%froude_cap_on = 1
#else
    cfg%froude_cap = 0.0_c_double; cfg%froude_cap_on = 0
#endif
    cfg%visc = Visc; cfg%fcor = fcor; cfg%wind_u = WindU; cfg%wind_v = WindV
    cfg%time_order = It_Order
    cfg%non_hydro = merge(1, 0, NON_HYDRO)
    cfg%hllc = merge(1, 0, ADV_HLLC)
    cfg%bc_x0 = Bc_X0; cfg%bc_xn = Bc_Xn; cfg%bc_y0 = Bc_Y0; cfg%bc_yn = Bc_Yn; cfg%bc_z0 = Bc_Z0
    cfg%itmax = itmax; cfg%tol = tol
    cfg%precision = 0
    cfg%solver = 0
    call get_environment_variable('NHW_SOLVER', sv, ln, st)
    if (st == 0 .and. sv(1:4) == 'line') cfg%solver = 1
    call nhw_shim_check(c_create(cfg, ctx), 'nhw_create')
    if (cfg%wavemaker == 5) then  ! the spectrum initialize.F read from irr2d.txt, packed
      call nhw_shim_check(c_set_internal(ctx, Xsource_West, Xsource_East, Ysource_Suth, Ysource_Nrth, &
                                         NumFreq, NumDir, Freq(1:NumFreq), Dire(1:NumDir), &
                                         reshape(Wave_Spc2d(1:NumDir, 1:NumFreq), [NumDir * NumFreq]), &
                                         reshape(Random_Phs(1:NumDir, 1:NumFreq), [NumDir * NumFreq])), &
                          'nhw_set_internal_wavemaker')
    endif
    call nhw_shim_check(c_set_grid(ctx, ad(dsig), ad(sig), ad(sigc)), 'nhw_set_grid')
    call all_fields(f)
    call nhw_shim_check(c_upload(ctx, f), 'nhw_upload')
    call nhw_shim_check(c_set_time(ctx, TIME, dt, RUN_STEP), 'nhw_set_time')
  end subroutine

  subroutine nhw_shim_mirror_time()
    use global, only: TIME, dt, RUN_STEP
    real(c_double) :: t_dev, dt_dev
    integer(c_int32_t) :: step_dev
    call nhw_shim_check(c_get_time(ctx, t_dev, dt_dev, step_dev), 'nhw_get_time')
    TIME = t_dev; dt = dt_dev; RUN_STEP = step_dev
  end subroutine

  logical function listed(env, step)
    character(len=*), intent(in) :: env
    integer, intent(in) :: step
    character(len=4096) :: steps
    integer :: ln, st, p0, p1, v, ios
    listed = .false.
    call get_environment_variable(env, steps, ln, st)
    if (st /= 0) return
    if (trim(steps) == 'all') then
      listed = .true.
      return
    endif
    p0 = 1
    do while (p0 <= ln)
      p1 = index(steps(p0:ln), ',')
      if (p1 == 0) then
        read(steps(p0:ln), *, iostat=ios) v
        p0 = ln + 1
      else
        read(steps(p0:p0+p1-2), *, iostat=ios) v
        p0 = p0 + p1
      endif
      if (ios == 0 .and. v == step) listed = .true.
    enddo
  end function

  subroutine snapshot()
    use global
    character(len=32) :: fname
    integer :: iu
    write(fname, '(A9,I5.5,A4)') 'nhw_snap_', RUN_STEP, '.bin'
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
  end subroutine

  ! after the last RK stage: bring the state to the host when the driver's
  ! output routines are about to read it (src/nhwave.F:268-296)
  subroutine nhw_shim_end_of_step()
    use global
    type(nhw_fields_t) :: f
    logical :: need, snap
    need = (Screen_Count + dt >= Screen_Intv)
    if (NSTAT > 0) need = need .or. (Plot_Count_Stat + dt >= Plot_Intv_Stat)
    if (TIME >= Plot_Start) need = need .or. (Plot_Count + dt >= Plot_Intv)
    snap = listed('NHW_SNAPSHOT_STEPS', RUN_STEP)
    if (need .or. snap) then
      call all_fields(f)
      call nhw_shim_check(c_download(ctx, f), 'nhw_download')
    elseif (OUT_M) then
      f%Eta = ad(Eta); f%Mask = ai(Mask)
      call nhw_shim_check(c_download(ctx, f), 'nhw_download')
    endif
    if (snap) call snapshot()
  end subroutine

end module nhw_shim

! --------------------------------------------------------------------------
! the substituted subroutines (same names and arguments as the reference)
! --------------------------------------------------------------------------
subroutine estimate_dt
  use global, only: RUN_STEP, dt, TIME
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_estimate_dt(ctx) bind(C, name='nhw_estimate_dt')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  integer(c_int) :: rc
  rc = c_estimate_dt(nhw_shim_ctx())
  if (rc == -3) then
    call nhw_shim_mirror_time()
    write(3,*) 'time step too small !!', dt
    stop
  endif
  call nhw_shim_check(rc, 'estimate_dt')
  call nhw_shim_mirror_time()
  write(3,*) RUN_STEP, dt, TIME
end subroutine estimate_dt

subroutine update_mask
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_update_mask(ctx) bind(C, name='nhw_update_mask')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_update_mask(nhw_shim_ctx()), 'update_mask')
end subroutine update_mask

subroutine update_wind
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_update_wind(ctx) bind(C, name='nhw_update_wind')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_update_wind(nhw_shim_ctx()), 'update_wind')
end subroutine update_wind

subroutine update_vars
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_update_vars(ctx) bind(C, name='nhw_update_vars')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_update_vars(nhw_shim_ctx()), 'update_vars')
end subroutine update_vars

subroutine source_terms
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_source_terms(ctx) bind(C, name='nhw_source_terms')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_source_terms(nhw_shim_ctx()), 'source_terms')
end subroutine source_terms

subroutine fluxes
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_fluxes(ctx) bind(C, name='nhw_fluxes')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_fluxes(nhw_shim_ctx()), 'fluxes')
end subroutine fluxes

subroutine eval_duvw(ISTEP)
  use global, only: It_Order
  use nhw_shim
  use iso_c_binding
  implicit none
  integer, intent(in) :: ISTEP
  interface
    integer(c_int) function c_eval_duvw(ctx, istep) bind(C, name='nhw_eval_duvw')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: istep
    end function
  end interface
  call nhw_shim_check(c_eval_duvw(nhw_shim_ctx(), int(ISTEP, c_int32_t)), 'eval_duvw')
  if (ISTEP == It_Order) call nhw_shim_end_of_step()
end subroutine eval_duvw

subroutine poisson_solver
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_poisson_solver(ctx) bind(C, name='nhw_poisson_solver')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_poisson_solver(nhw_shim_ctx()), 'poisson_solver')
end subroutine poisson_solver

subroutine update_wave_bc
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_update_wave_bc(ctx) bind(C, name='nhw_update_wave_bc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_update_wave_bc(nhw_shim_ctx()), 'update_wave_bc')
end subroutine update_wave_bc

subroutine sponge_damping
  use nhw_shim
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_sponge_damping(ctx) bind(C, name='nhw_sponge_damping')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_sponge_damping(nhw_shim_ctx()), 'sponge_damping')
end subroutine sponge_damping

subroutine eval_turb(ISTEP)
  use nhw_shim
  use iso_c_binding
  implicit none
  integer, intent(in) :: ISTEP
  interface
    integer(c_int) function c_eval_turb(ctx) bind(C, name='nhw_eval_turb')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
  call nhw_shim_check(c_eval_turb(nhw_shim_ctx()), 'eval_turb')
end subroutine eval_turb
