!! ========================================================================== !!
!! Fortran host driver over the MI355X spatial operator.
!!
!! Keeps the reference's game loop and hook order and replaces the per-quad
!! kernel/timedisc work with C-ABI calls:
!!
!!   hook_init            -> nemo_gpu_create, nemo_gpu_upload_mesh/state
!!                           (kernel_share_mod.f90:56-151, mesh_alloc, hook_init_810)
!!   mainloop             -> driver_mod.f90:62-150
!!     hook_calc_timestep -> nemo_gpu_calc_timestep (kernel_timestep_mod.f90:19-63)
!!     hook_evolve        -> timedisc_evolve_mod.f90:20-84: per stage
!!                           rktimestep = timestep*dt_coeffs(stage), then
!!                           nemo_gpu_stage (pre-stage faces + rkstep_cb)
!!     status check       -> nemo_gpu_synchronize (rt%stat%ok, :68-79)
!!     heartbeat          -> perf index (driver_mod.f90:140-149)
!!
!! Usage: nemo_gpu_driver SETUP N_ELEM N_POLY NSTEPS PREFIX
!!   SETUP  = vortex (isentropic vortex: weak form, Rusanov, SSPRK(5,4))
!!          | kh     (Kelvin-Helmholtz: EC flux differencing, HLLC, PP
!!                    blending, LSRK 5-4, positivity repair)
!!          | ot     (Orszag-Tang, GLM-MHD: EC flux differencing, HLLD,
!!                    PP blending, LSRK 5-4, positivity repair; the GLM
!!                    speed and damping follow each step's CFL reduction)
!!          | sod    (Sod's shock tube between slip walls across x, periodic
!!                    in y and z: the kh scheme with boundary faces)
!! Writes PREFIX.init.bin, PREFIX.final.bin (nemo_setup_mod::write_state)
!! and PREFIX.dt.txt (one time step per line).
!! ========================================================================== !!
program nemo_gpu_driver

use iso_c_binding
use nemo_gpu_mod
use nemo_setup_mod

implicit none

type :: runtime_t
    real(dp) :: simtime = 0.0_dp
    real(dp) :: timestep = 0.0_dp
    real(dp) :: maxtimestep = 0.0_dp
    real(dp) :: rktimestep = 0.0_dp
    real(dp) :: rksimtime = 0.0_dp
    integer  :: rkstage = 0
    integer  :: timesteps = 0
    logical  :: ok = .true.
    character(len=512) :: message = ''
end type

integer, parameter :: nstages = 5
!! dt_coeffs: low-storage/4-5/timedisc_const_mod.f90:29 (b_rk) and
!! ssp/4-5/timedisc_const_mod.f90:36
real(dp), parameter :: lsrk_dt_coeffs(nstages) = (/ 0.0_dp, &
    1432997174477.0_dp/9575080441755.0_dp, 2526269341429.0_dp/6820363962896.0_dp, &
    2006345519317.0_dp/3224310063776.0_dp, 2802321613138.0_dp/2924317926251.0_dp /)
real(dp), parameter :: ssp_dt_coeffs(nstages) = (/ 1.0_dp, 0.39175222700392_dp, &
    0.58607968896780_dp, 0.47454236302687_dp, 0.93501063100924_dp /)

type(runtime_t) :: rt
type(mesh_t) :: mesh
type(nemo_gpu_config) :: cfg
type(c_ptr) :: h
real(dp), allocatable :: state(:,:,:,:,:), xi(:), wi(:), dmat(:), vinv(:), dts(:)
integer(c_int32_t), allocatable :: bfaces(:,:), no_mortars(:)
integer(c_int64_t) :: nbound
character(len=1024) :: setup, prefix, arg
integer :: n, n_poly, nsteps, nn, nvars, unit, it
integer(c_int) :: rc
integer(c_int64_t) :: t0, t1, rate
real(dp) :: perfidx

call get_command_argument(1, setup)
call get_command_argument(2, arg); read(arg, *) n
call get_command_argument(3, arg); read(arg, *) n_poly
call get_command_argument(4, arg); read(arg, *) nsteps
call get_command_argument(5, prefix)

nn = n_poly + 1
cfg%n_poly = n_poly
cfg%equations = NEMO_EQ_EULER
cfg%precision = 0
cfg%gamma = 7.0_dp/5.0_dp
cfg%alpha_max = 0.5_dp
cfg%alpha_min = 0.001_dp
cfg%cfl = 0.5_dp
cfg%mu0 = 1.0_dp
select case (trim(setup))
case ('vortex')
    cfg%volume = NEMO_VOL_WEAK
    cfg%surface = NEMO_SURF_RUSANOV
    cfg%indicator = NEMO_IND_NONE
    cfg%timedisc = NEMO_TD_SSPRK54
    cfg%repair = 0
case ('kh')
    cfg%volume = NEMO_VOL_EC
    cfg%surface = NEMO_SURF_HLLC
    cfg%indicator = NEMO_IND_PP
    cfg%timedisc = NEMO_TD_LSRK54
    cfg%repair = 1
case ('sod')
    cfg%volume = NEMO_VOL_EC
    cfg%surface = NEMO_SURF_HLLC
    cfg%indicator = NEMO_IND_PP
    cfg%timedisc = NEMO_TD_LSRK54
    cfg%repair = 1
case ('ot')
    cfg%equations = NEMO_EQ_GLM_MHD
    cfg%gamma = 5.0_dp/3.0_dp
    cfg%volume = NEMO_VOL_EC
    cfg%surface = NEMO_SURF_HLLD
    cfg%indicator = NEMO_IND_PP
    cfg%timedisc = NEMO_TD_LSRK54
    cfg%repair = 1
case default
    write(*,*) 'unknown setup: ', trim(setup)
    stop 2
end select

call hook_init()
call write_state(trim(prefix)//'.init.bin', n_poly, mesh%nelem, state)

allocate(dts(nsteps))
call system_clock(t0, rate)
call mainloop()
call system_clock(t1)

if (.not. rt%ok) then
    write(*,'(a,a)') '[ERROR] ', trim(rt%message)
    rc = nemo_gpu_destroy(h)
    stop 1
end if

call check(nemo_gpu_download_state(h, state), 'nemo_gpu_download_state')
call write_state(trim(prefix)//'.final.bin', n_poly, mesh%nelem, state)
open(newunit=unit, file=trim(prefix)//'.dt.txt', status='replace')
do it = 1, rt%timesteps
    write(unit, '(es25.17)') dts(it)
end do
close(unit)

perfidx = real(t1 - t0, dp)/real(rate, dp) / &
          real(mesh%nelem*int(nn, c_int64_t)**3*rt%timesteps*nstages, dp)*1.0e6_dp
write(*,'(a,i8,2x,es12.4,2x,es12.4,4x,a,es12.4)') '[   FINISH]', rt%timesteps, rt%simtime, rt%maxtimestep, &
    'perf. index [us/DOF/stage] ', perfidx
rc = nemo_gpu_destroy(h)

contains

subroutine check(code, where)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: where
    if (code /= NEMO_OK) then
        rt%ok = .false.
        rt%message = trim(where)//': '//trim(nemo_gpu_last_error())
        write(*,'(a,a)') '[ERROR] ', trim(rt%message)
        stop 1
    end if
end subroutine

subroutine hook_init()
    allocate(xi(nn), wi(nn), dmat(nn*nn), vinv(nn*nn))
    call check(nemo_gpu_operators(int(n_poly, c_int32_t), xi, wi, dmat, vinv), 'nemo_gpu_operators')
    call check(nemo_gpu_create(cfg, h), 'nemo_gpu_create')
    call mesh_uniform(n, mesh)
    if (trim(setup) == 'sod') then
        call mesh_walls_x(mesh, NEMO_BC_SLIP_WALL, nbound, bfaces)
        allocate(no_mortars(6))
        call check(nemo_gpu_upload_mesh_bc(h, mesh%nelem, mesh%nbr, mesh%delta, 0_c_int64_t, no_mortars, &
                                           nbound, bfaces), 'nemo_gpu_upload_mesh_bc')
    else
        call check(nemo_gpu_upload_mesh(h, mesh%nelem, mesh%nbr, mesh%delta), 'nemo_gpu_upload_mesh')
    end if
    nvars = merge(9, 5, cfg%equations == NEMO_EQ_GLM_MHD)
    allocate(state(nn, nn, nn, nvars, mesh%nelem))
    select case (trim(setup))
    case ('vortex')
        call fill_vortex(mesh, nn, xi, cfg%gamma, state)
    case ('kh')
        call fill_kh(mesh, nn, xi, cfg%gamma, 0.01_dp, state)
    case ('sod')
        call fill_sod(mesh, nn, xi, cfg%gamma, state)
    case default
        call fill_orszag_tang(mesh, nn, xi, cfg%gamma, 1.0e-3_dp, state)
    end select
    call check(nemo_gpu_upload_state(h, state), 'nemo_gpu_upload_state')
end subroutine

subroutine hook_calc_timestep()
    real(dp) :: dt, lmax, vmax
    call check(nemo_gpu_calc_timestep(h, dt, lmax, vmax), 'nemo_gpu_calc_timestep')
    rt%maxtimestep = dt
end subroutine

subroutine hook_evolve()
    integer :: stage
    real(dp) :: coeff
    do stage = 1, nstages
        if (cfg%timedisc == NEMO_TD_LSRK54) then
            coeff = lsrk_dt_coeffs(stage)
        else
            coeff = ssp_dt_coeffs(stage)
        end if
        rt%rkstage = stage
        rt%rktimestep = rt%timestep*coeff
        rt%rksimtime = rt%simtime + rt%rktimestep
        call check(nemo_gpu_stage(h, int(stage, c_int32_t), rt%timestep), 'nemo_gpu_stage')
    end do
    call check(nemo_gpu_synchronize(h), 'nemo_gpu_synchronize')
end subroutine

subroutine mainloop()
    integer :: heartbeat
    heartbeat = 0
    do while (rt%ok .and. rt%timesteps < nsteps)
        call hook_calc_timestep()
        rt%timestep = rt%maxtimestep
        call hook_evolve()
        rt%simtime = rt%simtime + rt%timestep
        rt%timesteps = rt%timesteps + 1
        dts(rt%timesteps) = rt%timestep
        heartbeat = heartbeat + 1
        if (heartbeat >= 64) then
            write(*,'(a,i8,2x,es12.4,2x,es12.4)') '[HEARTBEAT]', rt%timesteps, rt%simtime, rt%maxtimestep
            heartbeat = 0
        end if
    end do
end subroutine

end program
