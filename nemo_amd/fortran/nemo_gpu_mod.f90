!! ========================================================================== !!
!! ISO_C_BINDING interfaces to the MI355X spatial operator (include/nemo_gpu.h).
!!
!! Modelled on the reference's own C bindings (mesh/p4est/wrapper/
!! p4wrap_interfaces_mod.f90:46-120): opaque type(c_ptr) handles, plain
!! arrays, integer status returns.  A Fortran host keeps nemo's driver loop
!! (driver_mod.f90:62-150) and calls these entry points from its kernel,
!! timedisc and timestep hooks.
!! ========================================================================== !!
module nemo_gpu_mod

use iso_c_binding

implicit none

integer(c_int), parameter :: NEMO_OK            = 0
integer(c_int), parameter :: NEMO_ERR_ARG       = 1
integer(c_int), parameter :: NEMO_ERR_HIP       = 2
integer(c_int), parameter :: NEMO_ERR_STATE     = 3
integer(c_int), parameter :: NEMO_ERR_NOT_READY = 4

integer(c_int32_t), parameter :: NEMO_EQ_EULER     = 0
integer(c_int32_t), parameter :: NEMO_EQ_GLM_MHD   = 1
integer(c_int32_t), parameter :: NEMO_BC_TRANSMISSIVE = 0
integer(c_int32_t), parameter :: NEMO_BC_SLIP_WALL    = 1
integer(c_int32_t), parameter :: NEMO_VOL_WEAK     = 0
integer(c_int32_t), parameter :: NEMO_VOL_CENTRAL  = 1
integer(c_int32_t), parameter :: NEMO_VOL_EC       = 2
integer(c_int32_t), parameter :: NEMO_SURF_RUSANOV = 0
integer(c_int32_t), parameter :: NEMO_SURF_HLLC    = 1
integer(c_int32_t), parameter :: NEMO_SURF_HLLD    = 2
integer(c_int32_t), parameter :: NEMO_IND_NONE     = 0
integer(c_int32_t), parameter :: NEMO_IND_PP       = 1
integer(c_int32_t), parameter :: NEMO_IND_MLP      = 2
integer(c_int32_t), parameter :: NEMO_MLP_NPARAMS  = 8705
integer(c_int32_t), parameter :: NEMO_ANALYZE_NSUMS = 16
integer(c_int32_t), parameter :: NEMO_TD_LSRK54    = 0
integer(c_int32_t), parameter :: NEMO_TD_SSPRK54   = 1

!! nemo_gpu_config: the reference's compile-time choices (preproc.py flags,
!! setup_config_mod constants) as a runtime record.
type, bind(c) :: nemo_gpu_config
    integer(c_int32_t) :: n_poly
    integer(c_int32_t) :: equations
    integer(c_int32_t) :: volume
    integer(c_int32_t) :: surface
    integer(c_int32_t) :: indicator
    integer(c_int32_t) :: timedisc
    integer(c_int32_t) :: repair
    integer(c_int32_t) :: precision
    real(c_double)     :: gamma
    real(c_double)     :: alpha_max
    real(c_double)     :: alpha_min
    real(c_double)     :: cfl
    real(c_double)     :: mu0
end type

interface

    function nemo_gpu_create(cfg, handle) result(rc) bind(c, name='nemo_gpu_create')
        import :: c_int, c_ptr, nemo_gpu_config
        type(nemo_gpu_config), intent(in) :: cfg
        type(c_ptr), intent(out)          :: handle
        integer(c_int)                    :: rc
    end function

    function nemo_gpu_destroy(handle) result(rc) bind(c, name='nemo_gpu_destroy')
        import :: c_int, c_ptr
        type(c_ptr), value :: handle
        integer(c_int)     :: rc
    end function

    function nemo_gpu_upload_mesh(handle, nelem, nbr, delta) result(rc) bind(c, name='nemo_gpu_upload_mesh')
        import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
        type(c_ptr), value                :: handle
        integer(c_int64_t), value         :: nelem
        integer(c_int32_t), intent(in)    :: nbr(*)
        real(c_double), intent(in)        :: delta(*)
        integer(c_int)                    :: rc
    end function

    !! adaptive mesh: sides with nquads = 4 as mortars(6, nmortar) =
    !! (big quad, its face, small quads 0..3), nbr = -(5m+1) / -(5m+2+k) on them
    function nemo_gpu_upload_mesh_amr(handle, nelem, nbr, delta, nmortar, mortars) result(rc) &
            bind(c, name='nemo_gpu_upload_mesh_amr')
        import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
        type(c_ptr), value                :: handle
        integer(c_int64_t), value         :: nelem, nmortar
        integer(c_int32_t), intent(in)    :: nbr(*), mortars(*)
        real(c_double), intent(in)        :: delta(*)
        integer(c_int)                    :: rc
    end function

    !! non-periodic mesh: boundary sides as bfaces(3, nbound) = (quad, its
    !! face, kind: NEMO_BC_TRANSMISSIVE or NEMO_BC_SLIP_WALL), with
    !! nbr = -(5*nmortar + b + 1) on boundary face b (0-based)
    function nemo_gpu_upload_mesh_bc(handle, nelem, nbr, delta, nmortar, mortars, nbound, bfaces) result(rc) &
            bind(c, name='nemo_gpu_upload_mesh_bc')
        import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
        type(c_ptr), value                :: handle
        integer(c_int64_t), value         :: nelem, nmortar, nbound
        integer(c_int32_t), intent(in)    :: nbr(*), mortars(*), bfaces(*)
        real(c_double), intent(in)        :: delta(*)
        integer(c_int)                    :: rc
    end function

    !! state transfer of mesh_adapt_wrapped: units(3, nunits) = (kind, first
    !! old quad, first new quad), kind 0 copy, 1 refine, 2 coarsen (0-based)
    function nemo_gpu_remesh(handle, nelem, nbr, delta, nmortar, mortars, nunits, units, fallbacks) result(rc) &
            bind(c, name='nemo_gpu_remesh')
        import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
        type(c_ptr), value                :: handle
        integer(c_int64_t), value         :: nelem, nmortar, nunits
        integer(c_int32_t), intent(in)    :: nbr(*), mortars(*), units(*)
        real(c_double), intent(in)        :: delta(*)
        integer(c_int64_t), intent(out)   :: fallbacks
        integer(c_int)                    :: rc
    end function

    !! volume integrals of hook_analyze: sums(NEMO_ANALYZE_NSUMS)
    function nemo_gpu_analyze(handle, sums) result(rc) bind(c, name='nemo_gpu_analyze')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value                :: handle
        real(c_double), intent(out)       :: sums(*)
        integer(c_int)                    :: rc
    end function

    !! allocation-free form: sums_dev is a device buffer of NEMO_ANALYZE_NSUMS
    !! doubles, filled asynchronously on the handle's stream
    function nemo_gpu_analyze_noalloc(handle, sums_dev) result(rc) bind(c, name='nemo_gpu_analyze_noalloc')
        import :: c_int, c_ptr
        type(c_ptr), value                :: handle, sums_dev
        integer(c_int)                    :: rc
    end function

    function nemo_gpu_upload_state(handle, u) result(rc) bind(c, name='nemo_gpu_upload_state')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value         :: handle
        real(c_double), intent(in) :: u(*)
        integer(c_int)             :: rc
    end function

    function nemo_gpu_download_state(handle, u) result(rc) bind(c, name='nemo_gpu_download_state')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value          :: handle
        real(c_double), intent(out) :: u(*)
        integer(c_int)              :: rc
    end function

    function nemo_gpu_rhs(handle, rhs) result(rc) bind(c, name='nemo_gpu_rhs')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value          :: handle
        real(c_double), intent(out) :: rhs(*)
        integer(c_int)              :: rc
    end function

    function nemo_gpu_indicator(handle, alpha) result(rc) bind(c, name='nemo_gpu_indicator')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value          :: handle
        real(c_double), intent(out) :: alpha(*)
        integer(c_int)              :: rc
    end function

    function nemo_gpu_calc_timestep(handle, dt, lmax, vmax) result(rc) bind(c, name='nemo_gpu_calc_timestep')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value          :: handle
        real(c_double), intent(out) :: dt, lmax, vmax
        integer(c_int)              :: rc
    end function

    !! equations_share_mod::glm_ch / glm_alpha (GLM-MHD)
    function nemo_gpu_set_glm(handle, ch, alpha) result(rc) bind(c, name='nemo_gpu_set_glm')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value    :: handle
        real(c_double), value :: ch, alpha
        integer(c_int)        :: rc
    end function

    function nemo_gpu_get_glm(handle, ch, alpha) result(rc) bind(c, name='nemo_gpu_get_glm')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value          :: handle
        real(c_double), intent(out) :: ch, alpha
        integer(c_int)              :: rc
    end function

    !! learned troubled-cell indicator
    function nemo_gpu_mlp_set_params(handle, params) result(rc) bind(c, name='nemo_gpu_mlp_set_params')
        import :: c_int, c_ptr, c_float
        type(c_ptr), value         :: handle
        real(c_float), intent(in)  :: params(*)
        integer(c_int)             :: rc
    end function

    function nemo_gpu_mlp_get_params(handle, params) result(rc) bind(c, name='nemo_gpu_mlp_get_params')
        import :: c_int, c_ptr, c_float
        type(c_ptr), value         :: handle
        real(c_float), intent(out) :: params(*)
        integer(c_int)             :: rc
    end function

    function nemo_gpu_mlp_set_policy(handle, sample_fraction, tolerance) result(rc) &
            bind(c, name='nemo_gpu_mlp_set_policy')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value    :: handle
        real(c_double), value :: sample_fraction, tolerance
        integer(c_int)        :: rc
    end function

    function nemo_gpu_mlp_stats(handle, judged, overridden, fallback) result(rc) bind(c, name='nemo_gpu_mlp_stats')
        import :: c_int, c_ptr, c_int64_t
        type(c_ptr), value              :: handle
        integer(c_int64_t), intent(out) :: judged, overridden, fallback
        integer(c_int)                  :: rc
    end function

    function nemo_gpu_indicator_features(handle, features, alpha_pp) result(rc) &
            bind(c, name='nemo_gpu_indicator_features')
        import :: c_int, c_ptr, c_float, c_double
        type(c_ptr), value          :: handle
        real(c_float), intent(out)  :: features(*)
        real(c_double), intent(out) :: alpha_pp(*)
        integer(c_int)              :: rc
    end function

    function nemo_gpu_mlp_forward(handle, n, features, y) result(rc) bind(c, name='nemo_gpu_mlp_forward')
        import :: c_int, c_ptr, c_int64_t, c_float
        type(c_ptr), value         :: handle
        integer(c_int64_t), value  :: n
        real(c_float), intent(in)  :: features(*)
        real(c_float), intent(out) :: y(*)
        integer(c_int)             :: rc
    end function

    function nemo_gpu_mlp_grad(handle, n, features, target, loss, grad) result(rc) bind(c, name='nemo_gpu_mlp_grad')
        import :: c_int, c_ptr, c_int64_t, c_float
        type(c_ptr), value         :: handle
        integer(c_int64_t), value  :: n
        real(c_float), intent(in)  :: features(*), target(*)
        real(c_float), intent(out) :: loss, grad(*)
        integer(c_int)             :: rc
    end function

    function nemo_gpu_mlp_train(handle, n, features, target, epochs, batch, lr, seed, loss_per_epoch) result(rc) &
            bind(c, name='nemo_gpu_mlp_train')
        import :: c_int, c_ptr, c_int64_t, c_int32_t, c_float
        type(c_ptr), value         :: handle
        integer(c_int64_t), value  :: n
        real(c_float), intent(in)  :: features(*), target(*)
        integer(c_int32_t), value  :: epochs, batch
        real(c_float), value       :: lr
        integer(c_int64_t), value  :: seed
        real(c_float), intent(out) :: loss_per_epoch(*)
        integer(c_int)             :: rc
    end function

    function nemo_gpu_stage(handle, stage, dt) result(rc) bind(c, name='nemo_gpu_stage')
        import :: c_int, c_ptr, c_int32_t, c_double
        type(c_ptr), value        :: handle
        integer(c_int32_t), value :: stage
        real(c_double), value     :: dt
        integer(c_int)            :: rc
    end function

    function nemo_gpu_step(handle, dt) result(rc) bind(c, name='nemo_gpu_step')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value    :: handle
        real(c_double), value :: dt
        integer(c_int)        :: rc
    end function

    function nemo_gpu_run(handle, nsteps, dt_hist) result(rc) bind(c, name='nemo_gpu_run')
        import :: c_int, c_ptr, c_int32_t
        type(c_ptr), value        :: handle
        integer(c_int32_t), value :: nsteps
        type(c_ptr), value        :: dt_hist
        integer(c_int)            :: rc
    end function

    function nemo_gpu_synchronize(handle) result(rc) bind(c, name='nemo_gpu_synchronize')
        import :: c_int, c_ptr
        type(c_ptr), value :: handle
        integer(c_int)     :: rc
    end function

    function nemo_gpu_get_status(handle, code) result(rc) bind(c, name='nemo_gpu_get_status')
        import :: c_int, c_ptr, c_int32_t
        type(c_ptr), value              :: handle
        integer(c_int32_t), intent(out) :: code
        integer(c_int)                  :: rc
    end function

    function nemo_gpu_operators(n_poly, x, w, d, vinv) result(rc) bind(c, name='nemo_gpu_operators')
        import :: c_int, c_int32_t, c_double
        integer(c_int32_t), value   :: n_poly
        real(c_double), intent(out) :: x(*), w(*), d(*), vinv(*)
        integer(c_int)              :: rc
    end function

    function nemo_gpu_error_message(buf, len) result(rc) bind(c, name='nemo_gpu_error_message')
        import :: c_int, c_int32_t, c_char
        character(kind=c_char), intent(out) :: buf(*)
        integer(c_int32_t), value           :: len
        integer(c_int)                      :: rc
    end function

end interface

contains

!! Last error message as a Fortran string (for raise_error).
function nemo_gpu_last_error() result(msg)
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer :: i, rc
    rc = nemo_gpu_error_message(buf, 512_c_int32_t)
    msg = ''
    do i = 1, 512
        if (buf(i) == c_null_char) exit
        msg(i:i) = buf(i)
    end do
end function

end module
