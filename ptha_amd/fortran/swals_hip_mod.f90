! ISO_C_BINDING interfaces to the MI355X time-stepping backend (include/swals_hip.h).
!
! This is the Fortran side of the drop-in boundary: SWALS's domain_type
! (propagation/SWALS/src/shallow_water/domain_mod.f90) keeps its setup, I/O and API,
! and its evolve_one_step hands the per-step work to these functions, following the
! pattern of the reference's existing bind(C) layer
! (propagation/SWALS/src/raster/read_raster_mod.f90:87-144). Arrays are passed with
! c_loc of the reference's own contiguous allocatables: domain%U(nx, ny, 4),
! domain%distance_bottom_edge, ... A non-zero status maps to generic_stop
! (propagation/SWALS/src/util/stop_mod.f90:8-33); swh_check does that here.
module swals_hip_mod
    use iso_c_binding
    implicit none
    private

    integer(c_int32_t), parameter, public :: SWH_METHOD_RK2 = 1, SWH_METHOD_RK2N = 2, SWH_METHOD_MIDPOINT = 3, &
        SWH_METHOD_EULER = 4, SWH_METHOD_LINEAR = 5, SWH_METHOD_LLPNF = 6, SWH_METHOD_LEAPFROG_NONLINEAR = 8
    integer(c_int32_t), parameter, public :: SWH_BC_STAGE_TRANSMISSIVE_NORMAL_MOMENTUM = 0, SWH_BC_FLATHER = 1, &
        SWH_BC_TRANSMISSIVE = 2, SWH_BC_STAGE_RADIATION_MOMENTUM = 3, SWH_BC_STAGE_TRANSMISSIVE_MOMENTUM = 4
    integer(c_int32_t), parameter, public :: SWH_PREC_FP64 = 0, SWH_PREC_FP32K = 1
    integer(c_int32_t), parameter, public :: SWH_MAXQ_MAX_STAGE = 1, SWH_MAXQ_MAX_SPEED = 2, &
        SWH_MAXQ_MAX_FLUX = 4, SWH_MAXQ_ARRIVAL_TIME = 8, SWH_MAXQ_TIME_OF_MAX_STAGE = 16, SWH_MAXQ_MIN_STAGE = 32

    ! Mirrors struct swh_domain_desc field by field
    type, bind(C), public :: swh_domain_desc
        integer(c_int32_t) :: nx = 0, ny = 0
        integer(c_int32_t) :: timestepping_method = SWH_METHOD_LINEAR
        integer(c_int32_t) :: precision = SWH_PREC_FP64
        integer(c_int32_t) :: linear_solver_is_truely_linear = 1
        integer(c_int32_t) :: include_coriolis = 0
        integer(c_int32_t) :: spherical = 0
        integer(c_int32_t) :: friction_type = 0
        integer(c_int32_t) :: exterior_cells_width = 2
        integer(c_int32_t) :: boundary_exterior(4) = 1
        integer(c_int32_t) :: max_U_mask = 0
        integer(c_int32_t) :: max_U_update_frequency = 1
        real(c_double) :: msl_linear = 0.0_c_double
        real(c_double) :: cfl = -1.0_c_double
        real(c_double) :: theta = -1.0_c_double
        real(c_double) :: linear_friction_coeff = 0.0_c_double
        real(c_double) :: maximum_timestep = -1.0_c_double
        real(c_double) :: arrival_stage_threshold_above_msl_linear = 0.01_c_double
        type(c_ptr) :: distance_bottom_edge = c_null_ptr
        type(c_ptr) :: distance_left_edge = c_null_ptr
        type(c_ptr) :: area_cell_y = c_null_ptr
        type(c_ptr) :: coriolis = c_null_ptr
        type(c_ptr) :: coriolis_bottom_edge = c_null_ptr
        type(c_ptr) :: tanlat_on_radius_earth = c_null_ptr
    end type

    ! Mirrors struct swh_scalars
    type, bind(C), public :: swh_scalars
        real(c_double) :: time, max_dt, evolve_step_dt, dt_last_update
        real(c_double) :: boundary_flux_store(4)
        real(c_double) :: boundary_flux_evolve_integral, boundary_flux_time_integral
        real(c_double) :: negative_depth_fix_counter
        integer(c_int64_t) :: nsteps_advanced
    end type

    ! Mirrors struct swh_domain_stats (compute_domain_statistics_new's results)
    type, bind(C), public :: swh_domain_stats
        real(c_double) :: max_stage, max_speed, min_stage, min_speed
        real(c_double) :: energy_potential_on_rho, energy_kinetic_on_rho, energy_total_on_rho
    end type

    integer(c_int32_t), parameter, public :: SWH_SUPERPOSE_SERIES = 0, SWH_SUPERPOSE_MAX_STAGE = 1

    ! Mirrors struct swh_nest_box: one two_way_nesting_comms_type box (1-based indices)
    type, bind(C), public :: swh_nest_box
        integer(c_int32_t) :: inds(6) = 0
        integer(c_int32_t) :: neighbour_domain_index = 0, neighbour_comms_index = 0
        integer(c_int32_t) :: my_domain_is_finer = 0, equal_cell_ratios = 0
        integer(c_int32_t) :: my_domain_staggered_grid = 0, neighbour_domain_staggered_grid = 0
        integer(c_int32_t) :: use_wetdry_limiting = 1
        integer(c_int32_t) :: buffer_size = 0, flux_start = 0, time_index = 0
        integer(c_int32_t) :: flux_rows(4) = 0, flux_nvar = 0
        real(c_double) :: cell_ratios(2) = 1.0_c_double
    end type

    public :: swh_md_create, swh_md_destroy, swh_md_set_priority, swh_md_set_flux_correct_dry_cells, &
        swh_md_add_box, swh_md_finalize, swh_md_make_initial_conditions_consistent, swh_md_evolve_one_step, &
        swh_md_evolve_steps, swh_md_volumes, swh_md_synchronize, swh_md_statistics, swh_md_check_stability, &
        swh_md_graph_capture, swh_md_graph_launch, swh_md_graph_destroy

    public :: swh_domain_statistics, swh_domain_statistics_noalloc, swh_domain_check_stability, swh_snapshot, &
        swh_snapshot_download, swh_checkpoint_save, swh_checkpoint_restore, swh_evolve_record_gauges_schedule

    public :: swh_batch_create, swh_batch_destroy, swh_batch_upload_elevation, swh_batch_upload_manning_squared, &
        swh_batch_upload_source, swh_batch_download_source, swh_batch_set_nsteps, swh_batch_gauges_set, &
        swh_batch_gauges_set_schedule, swh_batch_evolve, swh_batch_gauges_download, swh_batch_status, &
        swh_batch_autotune, swh_batch_synchronize, swh_batch_superpose

    public :: swh_last_error, swh_domain_create, swh_domain_destroy, swh_upload_U, swh_download_U, &
        swh_upload_manning_squared, swh_set_time, swh_set_stage_boundary_sine, swh_set_flather_boundary, &
        swh_set_boundary, &
        swh_set_forcing_patch, swh_clear_forcing_patch, swh_evolve_one_step, swh_dispersion_enable, &
        swh_dispersion_store_last_U, swh_dispersion_append_extrapolation, swh_dispersion_solve, &
        swh_evolve_steps, swh_get_scalars, swh_volume_interior, swh_stationary_timestep_max, &
        swh_download_max_U, swh_gauges_set, swh_gauges_sample, swh_gauges_download, swh_synchronize, &
        swh_domain_set_launch, swh_domain_set_launch2, swh_upload_state, swh_check

    interface
        integer(c_int) function swh_last_error(buf, len) bind(C, name='swh_last_error')
            import c_int, c_char
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int), value :: len
        end function
        integer(c_int) function swh_domain_create(desc, h) bind(C, name='swh_domain_create')
            import c_int, c_ptr, swh_domain_desc
            type(swh_domain_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function swh_domain_destroy(h) bind(C, name='swh_domain_destroy')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function swh_upload_U(h, U) bind(C, name='swh_upload_U')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: U(*)
        end function
        ! A new initial condition on the same grid (elevation kept): c_loc of stage, UH,
        ! VH arrays of (nx, ny), or c_null_ptr for zeros.
        integer(c_int) function swh_upload_state(h, stage, uh, vh) bind(C, name='swh_upload_state')
            import c_int, c_ptr
            type(c_ptr), value :: h, stage, uh, vh
        end function
        integer(c_int) function swh_download_U(h, U) bind(C, name='swh_download_U')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: U(*)
        end function
        integer(c_int) function swh_upload_manning_squared(h, m) bind(C, name='swh_upload_manning_squared')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: m(*)
        end function
        integer(c_int) function swh_set_time(h, t) bind(C, name='swh_set_time')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: t
        end function
        integer(c_int) function swh_set_stage_boundary_sine(h, amp, period, msl, sides) &
                bind(C, name='swh_set_stage_boundary_sine')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            real(c_double), value :: amp, period, msl
            integer(c_int32_t), intent(in) :: sides(4)
        end function
        ! forcing_patch_type with apply_forcing_patch (forcing_mod.f90:24-56, 80-156); the
        ! patch's work array as the reference holds it, work(i0:i1, j0:j1, k0:k1)
        integer(c_int) function swh_set_forcing_patch(h, start_time, end_time, i0, i1, j0, j1, k0, k1, work) &
                bind(C, name='swh_set_forcing_patch')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            real(c_double), value :: start_time, end_time
            integer(c_int32_t), value :: i0, i1, j0, j1, k0, k1
            real(c_double), intent(in) :: work(*)
        end function
        integer(c_int) function swh_clear_forcing_patch(h) bind(C, name='swh_clear_forcing_patch')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        ! any boundary subroutine of boundary_mod.f90 (SWH_BC_*) on sides [N, E, S, W]
        integer(c_int) function swh_set_boundary(h, kind, sides, use_stage_function, amp, period, msl) &
                bind(C, name='swh_set_boundary')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: kind
            integer(c_int32_t), intent(in) :: sides(4)
            integer(c_int32_t), value :: use_stage_function
            real(c_double), value :: amp, period, msl
        end function
        ! flather_boundary (boundary_mod.f90:502-720) on sides [N, E, S, W]
        integer(c_int) function swh_set_flather_boundary(h, sides, use_stage_function, amp, period, msl) &
                bind(C, name='swh_set_flather_boundary')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), intent(in) :: sides(4)
            integer(c_int32_t), value :: use_stage_function
            real(c_double), value :: amp, period, msl
        end function
        ! dispersion (linear_dispersive_solver_mod.f90), staggered methods
        integer(c_int) function swh_dispersion_enable(h, dlon, dlat, td1, td2, niter) &
                bind(C, name='swh_dispersion_enable')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            real(c_double), value :: dlon, dlat, td1, td2
            integer(c_int32_t), value :: niter
        end function
        integer(c_int) function swh_dispersion_store_last_U(h) bind(C, name='swh_dispersion_store_last_U')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function swh_dispersion_append_extrapolation(h) &
                bind(C, name='swh_dispersion_append_extrapolation')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function swh_dispersion_solve(h, rhs_is_up_to_date, estimate_forward, forward_time) &
                bind(C, name='swh_dispersion_solve')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: rhs_is_up_to_date, estimate_forward
            real(c_double), value :: forward_time
        end function
        integer(c_int) function swh_evolve_one_step(h, timestep) bind(C, name='swh_evolve_one_step')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: timestep
        end function
        integer(c_int) function swh_evolve_steps(h, timestep, nsteps) bind(C, name='swh_evolve_steps')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            real(c_double), value :: timestep
            integer(c_int32_t), value :: nsteps
        end function
        integer(c_int) function swh_get_scalars(h, s) bind(C, name='swh_get_scalars')
            import c_int, c_ptr, swh_scalars
            type(c_ptr), value :: h
            type(swh_scalars), intent(out) :: s
        end function
        integer(c_int) function swh_volume_interior(h, v) bind(C, name='swh_volume_interior')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function swh_stationary_timestep_max(h, v) bind(C, name='swh_stationary_timestep_max')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function swh_download_max_U(h, which, out) bind(C, name='swh_download_max_U')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: which
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function swh_gauges_set(h, n, ij, max_samples) bind(C, name='swh_gauges_set')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: n, max_samples
            integer(c_int32_t), intent(in) :: ij(*)
        end function
        integer(c_int) function swh_gauges_sample(h) bind(C, name='swh_gauges_sample')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function swh_gauges_download(h, values, times, nsamples) bind(C, name='swh_gauges_download')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            real(c_double), intent(out) :: values(*), times(*)
            integer(c_int32_t), intent(out) :: nsamples
        end function
        integer(c_int) function swh_synchronize(h) bind(C, name='swh_synchronize')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function swh_domain_set_launch(h, block_x, rows_per_block, variant) &
                bind(C, name='swh_domain_set_launch')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: block_x, rows_per_block, variant
        end function
        integer(c_int) function swh_domain_set_launch2(h, block_x, rows_per_block) &
                bind(C, name='swh_domain_set_launch2')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: block_x, rows_per_block
        end function
        ! ---- multidomain (one GPU) ----
        integer(c_int) function swh_md_create(n, domains, all_dx, all_methods, trf, md) bind(C, name='swh_md_create')
            import c_int, c_ptr, c_int32_t, c_double
            integer(c_int32_t), value :: n
            type(c_ptr), intent(in) :: domains(*)
            real(c_double), intent(in) :: all_dx(*)
            integer(c_int32_t), intent(in) :: all_methods(*), trf(*)
            type(c_ptr), intent(out) :: md
        end function
        integer(c_int) function swh_md_destroy(md) bind(C, name='swh_md_destroy')
            import c_int, c_ptr
            type(c_ptr), value :: md
        end function
        integer(c_int) function swh_md_set_priority(md, k, pdi, pnp) bind(C, name='swh_md_set_priority')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: md
            integer(c_int32_t), value :: k
            integer(c_int32_t), intent(in) :: pdi(*), pnp(*)
        end function
        integer(c_int) function swh_md_set_flux_correct_dry_cells(md, k, on) &
                bind(C, name='swh_md_set_flux_correct_dry_cells')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: md
            integer(c_int32_t), value :: k, on
        end function
        integer(c_int) function swh_md_add_box(md, k, is_recv, box) bind(C, name='swh_md_add_box')
            import c_int, c_ptr, c_int32_t, swh_nest_box
            type(c_ptr), value :: md
            integer(c_int32_t), value :: k, is_recv
            type(swh_nest_box), intent(in) :: box
        end function
        integer(c_int) function swh_md_finalize(md) bind(C, name='swh_md_finalize')
            import c_int, c_ptr
            type(c_ptr), value :: md
        end function
        integer(c_int) function swh_md_make_initial_conditions_consistent(md) &
                bind(C, name='swh_md_make_initial_conditions_consistent')
            import c_int, c_ptr
            type(c_ptr), value :: md
        end function
        integer(c_int) function swh_md_evolve_one_step(md, dt) bind(C, name='swh_md_evolve_one_step')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: md
            real(c_double), value :: dt
        end function
        integer(c_int) function swh_md_evolve_steps(md, dt, nsteps) bind(C, name='swh_md_evolve_steps')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: md
            real(c_double), value :: dt
            integer(c_int32_t), value :: nsteps
        end function
        integer(c_int) function swh_md_volumes(md, out) bind(C, name='swh_md_volumes')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: md
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function swh_md_synchronize(md) bind(C, name='swh_md_synchronize')
            import c_int, c_ptr
            type(c_ptr), value :: md
        end function
        ! print_multidomain's statistics (multidomain_mod.f90:2301-2447) and the stability scan
        integer(c_int) function swh_md_statistics(md, per_domain, global, energy_is_finite) &
                bind(C, name='swh_md_statistics')
            import c_int, c_ptr, c_int32_t, swh_domain_stats
            type(c_ptr), value :: md
            type(swh_domain_stats), intent(out) :: per_domain(*)
            type(swh_domain_stats), intent(out) :: global
            integer(c_int32_t), intent(out) :: energy_is_finite
        end function
        integer(c_int) function swh_md_check_stability(md, n_bad) bind(C, name='swh_md_check_stability')
            import c_int, c_ptr, c_int64_t
            type(c_ptr), value :: md
            integer(c_int64_t), intent(out) :: n_bad
        end function
        ! The global step as HIP graphs: capture once, replay nsteps global steps
        integer(c_int) function swh_md_graph_capture(md, dt) bind(C, name='swh_md_graph_capture')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: md
            real(c_double), value :: dt
        end function
        integer(c_int) function swh_md_graph_launch(md, nsteps) bind(C, name='swh_md_graph_launch')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: md
            integer(c_int32_t), value :: nsteps
        end function
        integer(c_int) function swh_md_graph_destroy(md) bind(C, name='swh_md_graph_destroy')
            import c_int, c_ptr
            type(c_ptr), value :: md
        end function
        ! compute_domain_statistics_new (domain_mod.f90:683-849) and the stability scan
        integer(c_int) function swh_domain_statistics(h, out) bind(C, name='swh_domain_statistics')
            import c_int, c_ptr, swh_domain_stats
            type(c_ptr), value :: h
            type(swh_domain_stats), intent(out) :: out
        end function
        integer(c_int) function swh_domain_statistics_noalloc(h, out) bind(C, name='swh_domain_statistics_noalloc')
            import c_int, c_ptr, swh_domain_stats
            type(c_ptr), value :: h
            type(swh_domain_stats), intent(out) :: out
        end function
        integer(c_int) function swh_domain_check_stability(h, n_bad) bind(C, name='swh_domain_check_stability')
            import c_int, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: n_bad
        end function
        ! Output staging: pack U(nx, ny, 4) now, download it later on a copy stream
        integer(c_int) function swh_snapshot(h) bind(C, name='swh_snapshot')
            import c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function swh_snapshot_download(h, U, time) bind(C, name='swh_snapshot_download')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: U(*)
            real(c_double), intent(out) :: time
        end function
        integer(c_int) function swh_checkpoint_save(h, slot) bind(C, name='swh_checkpoint_save')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: slot
        end function
        integer(c_int) function swh_checkpoint_restore(h, slot) bind(C, name='swh_checkpoint_restore')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: slot
        end function
        ! gauges written at the listed steps (the driver's write-time rule, generic_model.f90:380-405)
        integer(c_int) function swh_evolve_record_gauges_schedule(h, timestep, nsamples, sample_steps) &
                bind(C, name='swh_evolve_record_gauges_schedule')
            import c_int, c_ptr, c_double, c_int32_t, c_int64_t
            type(c_ptr), value :: h
            real(c_double), value :: timestep
            integer(c_int32_t), value :: nsamples
            integer(c_int64_t), intent(in) :: sample_steps(*)
        end function
        ! Batched unit sources on one shared grid (one run_16.sh model run per source)
        integer(c_int) function swh_batch_create(desc, nsources, b) bind(C, name='swh_batch_create')
            import c_int, c_ptr, c_int32_t, swh_domain_desc
            type(swh_domain_desc), intent(in) :: desc
            integer(c_int32_t), value :: nsources
            type(c_ptr), intent(out) :: b
        end function
        integer(c_int) function swh_batch_destroy(b) bind(C, name='swh_batch_destroy')
            import c_int, c_ptr
            type(c_ptr), value :: b
        end function
        integer(c_int) function swh_batch_upload_elevation(b, elev) bind(C, name='swh_batch_upload_elevation')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: b
            real(c_double), intent(in) :: elev(*)
        end function
        integer(c_int) function swh_batch_upload_manning_squared(b, m) bind(C, name='swh_batch_upload_manning_squared')
            import c_int, c_ptr, c_double
            type(c_ptr), value :: b
            real(c_double), intent(in) :: m(*)
        end function
        integer(c_int) function swh_batch_upload_source(b, k, U3) bind(C, name='swh_batch_upload_source')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: k
            real(c_double), intent(in) :: U3(*)
        end function
        integer(c_int) function swh_batch_download_source(b, k, U3) bind(C, name='swh_batch_download_source')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: k
            real(c_double), intent(out) :: U3(*)
        end function
        integer(c_int) function swh_batch_set_nsteps(b, nsteps_per_source) bind(C, name='swh_batch_set_nsteps')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), intent(in) :: nsteps_per_source(*)
        end function
        integer(c_int) function swh_batch_gauges_set(b, n, ij, every, max_samples) bind(C, name='swh_batch_gauges_set')
            import c_int, c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: n, every, max_samples
            integer(c_int32_t), intent(in) :: ij(*)
        end function
        integer(c_int) function swh_batch_gauges_set_schedule(b, n, ij, nsamples, sample_steps) &
                bind(C, name='swh_batch_gauges_set_schedule')
            import c_int, c_ptr, c_int32_t, c_int64_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: n, nsamples
            integer(c_int32_t), intent(in) :: ij(*)
            integer(c_int64_t), intent(in) :: sample_steps(*)
        end function
        integer(c_int) function swh_batch_evolve(b, dt, nsteps) bind(C, name='swh_batch_evolve')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: b
            real(c_double), value :: dt
            integer(c_int32_t), value :: nsteps
        end function
        integer(c_int) function swh_batch_gauges_download(b, k, values, nsamples) bind(C, name='swh_batch_gauges_download')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: k
            real(c_double), intent(out) :: values(*)
            integer(c_int32_t), intent(out) :: nsamples
        end function
        integer(c_int) function swh_batch_status(b, step, n_active) bind(C, name='swh_batch_status')
            import c_int, c_ptr, c_int32_t, c_int64_t
            type(c_ptr), value :: b
            integer(c_int64_t), intent(out) :: step
            integer(c_int32_t), intent(out) :: n_active
        end function
        integer(c_int) function swh_batch_autotune(b, dt, reps, best, best_ms) bind(C, name='swh_batch_autotune')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: b
            real(c_double), value :: dt
            integer(c_int32_t), value :: reps
            integer(c_int32_t), intent(out) :: best(3)
            real(c_double), intent(out) :: best_ms
        end function
        integer(c_int) function swh_batch_synchronize(b) bind(C, name='swh_batch_synchronize')
            import c_int, c_ptr
            type(c_ptr), value :: b
        end function
        ! make_tsunami_event_from_unit_sources (sum_tsunami_unit_sources.R:555-711) on the device
        integer(c_int) function swh_batch_superpose(b, n_events, event_ptr, event_src, event_w, event_slip, msl, &
                summary, out, nvalid, kernel_ms) bind(C, name='swh_batch_superpose')
            import c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: n_events, summary
            integer(c_int32_t), intent(in) :: event_ptr(*), event_src(*)
            type(c_ptr), value :: event_w, event_slip
            real(c_double), value :: msl
            real(c_double), intent(out) :: out(*)
            integer(c_int32_t), intent(out) :: nvalid(*)
            type(c_ptr), value :: kernel_ms
        end function
    end interface

contains

    subroutine swh_check(status, what)
        !! Stops with the backend's message on a non-zero status (the reference's
        !! generic_stop behaviour: print, then stop).
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: what
        character(kind=c_char) :: buf(512)
        character(len=512) :: msg
        integer :: k, rc
        if(status == 0) return
        buf = c_null_char
        rc = swh_last_error(buf, 512_c_int)
        msg = ''
        do k = 1, 512
            if(buf(k) == c_null_char) exit
            msg(k:k) = buf(k)
        end do
        write(*, *) 'swals_hip: ', trim(what), ' failed (status ', status, '): ', trim(msg)
        error stop 1
    end subroutine

end module
