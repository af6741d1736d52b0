! Test infrastructure only: a bind(C) shim over the reference SWALS domain_type,
! compiled together with the reference's own Fortran modules (read in place from the
! reference tree by oracle/Makefile) into oracle/_ref/libswals_ref_{cart,sph}.so.
!
! It lets the CPU test-suite and the bench's CPU-baseline leg drive the *real*
! reference solver (domain%evolve_one_step, propagation/SWALS/src/shallow_water/
! domain_mod.f90:1679-1813) from Python through ctypes. Nothing in the product links it.
!
! Handles are small integers indexing a fixed pool of domain_type objects.
module swals_ref_capi
    use iso_c_binding
    use global_mod, only: dp, ip, charlen, force_double
    use domain_mod, only: domain_type, STG, UH, VH, ELV
    use boundary_mod, only: boundary_stage_transmissive_normal_momentum, flather_boundary, transmissive_boundary, &
        boundary_stage_radiation_momentum, boundary_stage_transmissive_momentum
    use forcing_mod, only: forcing_patch_type, apply_forcing_patch
    use multidomain_mod, only: multidomain_type
    use timestepping_metadata_mod, only: timestepping_method_index
    implicit none

    integer, parameter :: max_domains = 32
    type(domain_type), target, save :: pool(max_domains)
    logical, save :: in_use(max_domains) = .false.

    ! Parameters of the sinusoidal stage boundary (one set per handle)
    real(dp), save :: bc_amp(max_domains) = 0.0_dp, bc_period(max_domains) = 1.0_dp, &
        bc_msl(max_domains) = 0.0_dp

    ! One forcing patch per handle (forcing_mod.f90:24-56)
    type(forcing_patch_type), target, save :: fpatch(max_domains)

    ! One reference multidomain (multidomain_mod.f90:171-276) for nesting checks
    type(multidomain_type), target, save :: md

contains

    function method_name(method) result(name)
        ! Method ids follow the order of the reference's timestepping_metadata table
        ! (propagation/SWALS/src/shallow_water/timestepping_metadata_mod.f90:69-140).
        integer(c_int), intent(in) :: method
        character(len=charlen) :: name
        select case(method)
        case(1); name = 'rk2'
        case(2); name = 'rk2n'
        case(3); name = 'midpoint'
        case(4); name = 'euler'
        case(5); name = 'linear'
        case(6); name = 'leapfrog_linear_plus_nonlinear_friction'
        case(8); name = 'leapfrog_nonlinear'
        case default; name = 'unknown'
        end select
    end function

    function ref_domain_create(nx, ny, lw, ll, method, truely_linear, msl_linear, &
            linear_friction_coeff, friction_type, cfl, theta) result(h) bind(C, name='ref_domain_create')
        integer(c_int), value :: nx, ny, method, truely_linear, friction_type
        real(c_double), intent(in) :: lw(2), ll(2)
        real(c_double), value :: msl_linear, linear_friction_coeff, cfl, theta
        integer(c_int) :: h
        integer :: k

        h = -1
        do k = 1, max_domains
            if(.not. in_use(k)) then
                h = k
                exit
            end if
        end do
        if(h < 0) return
        in_use(h) = .true.

        associate(d => pool(h))
            d%timestepping_method = method_name(method)
            d%linear_solver_is_truely_linear = (truely_linear /= 0)
            d%msl_linear = msl_linear
            d%linear_friction_coeff = linear_friction_coeff
            if(friction_type == 1) then
                d%friction_type = 'chezy'
            else
                d%friction_type = 'manning'
            end if
            if(cfl > 0.0_c_double) d%cfl = cfl
            if(theta > 0.0_c_double) d%theta = theta
            d%record_max_U = .true.
            d%nontemporal_grids_to_store = [character(len=charlen):: &
                'max_stage', 'max_speed', 'max_flux', 'arrival_time', 'time_of_max_stage', 'min_stage']
            call d%allocate_quantities(real(lw, dp), [nx, ny], real(ll, dp), &
                create_output_files=.false., verbose=.false.)
        end associate
    end function

    subroutine ref_domain_destroy(h) bind(C, name='ref_domain_destroy')
        integer(c_int), value :: h
        type(domain_type) :: fresh
        if(h < 1 .or. h > max_domains) return
        pool(h) = fresh
        in_use(h) = .false.
    end subroutine

    subroutine ref_domain_set_U(h, U) bind(C, name='ref_domain_set_U')
        integer(c_int), value :: h
        real(c_double), intent(in) :: U(*)
        integer :: n
        n = pool(h)%nx(1) * pool(h)%nx(2)
        pool(h)%U = reshape(real(U(1:4*n), dp), [pool(h)%nx(1), pool(h)%nx(2), 4])
    end subroutine

    subroutine ref_domain_get_U(h, U) bind(C, name='ref_domain_get_U')
        integer(c_int), value :: h
        real(c_double), intent(out) :: U(*)
        integer :: n
        n = pool(h)%nx(1) * pool(h)%nx(2)
        U(1:4*n) = real(reshape(pool(h)%U(:,:,1:4), [4*n]), c_double)
    end subroutine

    subroutine ref_domain_set_forcing(h, start_time, end_time, i0, i1, j0, j1, k0, k1, work) &
            bind(C, name='ref_domain_set_forcing')
        ! A forcing patch applied by the reference's own apply_forcing_patch
        ! (forcing_mod.f90:158-239): work(i0:i1, j0:j1, k0:k1) in Fortran order.
        integer(c_int), value :: h, i0, i1, j0, j1, k0, k1
        real(c_double), value :: start_time, end_time
        real(c_double), intent(in) :: work(*)
        integer :: n
        call fpatch(h)%setup(real(start_time, dp), real(end_time, dp), i0, i1, j0, j1, k0, k1)
        n = (i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1)
        fpatch(h)%forcing_work = reshape(real(work(1:n), dp), [i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1])
        pool(h)%forcing_context_cptr = c_loc(fpatch(h))
        pool(h)%forcing_subroutine => apply_forcing_patch
    end subroutine

    subroutine ref_domain_sample(h, n, ij, out) bind(C, name='ref_domain_sample')
        ! [STG, UH, VH] at n cells (1-based i, j pairs): out((var - 1) * n + g)
        integer(c_int), value :: h, n
        integer(c_int), intent(in) :: ij(2, *)
        real(c_double), intent(out) :: out(*)
        integer :: g, v
        do v = 1, 3
            do g = 1, n
                out((v - 1) * n + g) = real(pool(h)%U(ij(1, g), ij(2, g), v), c_double)
            end do
        end do
    end subroutine

    subroutine ref_domain_set_manning_squared(h, m) bind(C, name='ref_domain_set_manning_squared')
        integer(c_int), value :: h
        real(c_double), intent(in) :: m(*)
        integer :: n
        n = pool(h)%nx(1) * pool(h)%nx(2)
        pool(h)%manning_squared = reshape(real(m(1:n), dp), [pool(h)%nx(1), pool(h)%nx(2)])
    end subroutine

    subroutine ref_domain_get_metrics(h, x, y, dbe, dle, area, cor, cor_be, tanlat) &
            bind(C, name='ref_domain_get_metrics')
        ! Coriolis and tan(lat)/R exist only in the spherical build; zeros otherwise.
        integer(c_int), value :: h
        real(c_double), intent(out) :: x(*), y(*), dbe(*), dle(*), area(*), cor(*), cor_be(*), tanlat(*)
        integer :: nx, ny
        nx = pool(h)%nx(1); ny = pool(h)%nx(2)
        x(1:nx) = pool(h)%x
        y(1:ny) = pool(h)%y
        dbe(1:(ny+1)) = pool(h)%distance_bottom_edge
        dle(1:(nx+1)) = pool(h)%distance_left_edge
        area(1:ny) = pool(h)%area_cell_y
        cor(1:ny) = 0.0_c_double; cor_be(1:(ny+1)) = 0.0_c_double; tanlat(1:ny) = 0.0_c_double
        if(allocated(pool(h)%coriolis)) cor(1:ny) = pool(h)%coriolis
        if(allocated(pool(h)%coriolis_bottom_edge)) cor_be(1:(ny+1)) = pool(h)%coriolis_bottom_edge
        if(allocated(pool(h)%tanlat_on_radius_earth)) tanlat(1:ny) = pool(h)%tanlat_on_radius_earth
    end subroutine

    function ref_domain_evolve(h, dt, nsteps) result(t) bind(C, name='ref_domain_evolve')
        ! nsteps calls of domain%evolve_one_step; dt <= 0 means "no timestep argument"
        ! (adaptive CFL step for the finite-volume methods). Returns domain%time.
        integer(c_int), value :: h, nsteps
        real(c_double), value :: dt
        real(c_double) :: t
        integer :: k
        do k = 1, nsteps
            if(dt > 0.0_c_double) then
                call pool(h)%evolve_one_step(real(dt, dp))
            else
                call pool(h)%evolve_one_step()
            end if
        end do
        t = pool(h)%time
    end function

    subroutine ref_domain_get_scalars(h, out) bind(C, name='ref_domain_get_scalars')
        ! out(1:12) = time, max_dt, evolve_step_dt, boundary_flux_time_integral,
        !             boundary_flux_evolve_integral, boundary_flux_store(1:4),
        !             negative_depth_fix_counter, nsteps_advanced, dt_last_update
        integer(c_int), value :: h
        real(c_double), intent(out) :: out(*)
        out(1) = pool(h)%time
        out(2) = pool(h)%max_dt
        out(3) = pool(h)%evolve_step_dt
        out(4) = pool(h)%boundary_flux_time_integral
        out(5) = pool(h)%boundary_flux_evolve_integral
        out(6:9) = pool(h)%boundary_flux_store(1:4)
        out(10) = real(pool(h)%negative_depth_fix_counter, c_double)
        out(11) = real(pool(h)%nsteps_advanced, c_double)
        out(12) = pool(h)%dt_last_update
    end subroutine

    subroutine ref_domain_set_record_max_U(h, flag) bind(C, name='ref_domain_set_record_max_U')
        ! Switches update_max_quantities on or off (domain_mod.f90:2252 tests record_max_U)
        integer(c_int), value :: h, flag
        pool(h)%record_max_U = (flag /= 0)
    end subroutine

    ! Dispersion (linear_dispersive_solver_mod.f90) driven as multidomain_mod.f90:1245-1410
    ! drives it for one domain: the caller appends the extrapolation level, stores U,
    ! steps, then solves.
    subroutine ref_domain_dispersion_setup(h, td1, td2, niter) bind(C, name='ref_domain_dispersion_setup')
        integer(c_int), value :: h, niter
        real(c_double), value :: td1, td2
        integer :: nx, ny
        nx = pool(h)%nx(1); ny = pool(h)%nx(2)
        pool(h)%use_dispersion = .true.
        ! the allocations of setup_dispersive_solver (linear_dispersive_solver_mod.f90:144-179)
        ! made here: its OpenMP first-touch loop takes its bound from a dummy argument that
        ! DEFAULT(PRIVATE) makes private, which faults when compiled with amdflang
        pool(h)%ds%is_staggered_grid = pool(h)%is_staggered_grid
        allocate(pool(h)%ds%RHS(nx, ny, UH:VH), pool(h)%ds%last_U(nx, ny, UH:ELV), pool(h)%ds%Ax(nx, ny, UH:VH))
        pool(h)%ds%RHS = 0.0_dp
        pool(h)%ds%last_U = 0.0_dp
        pool(h)%ds%Ax = 0.0_dp
        call pool(h)%ds%qet%setup(nx, ny, 2)
        pool(h)%ds%td1 = td1
        pool(h)%ds%td2 = td2
        pool(h)%ds%tridiagonal_inner_iter = niter
    end subroutine

    subroutine ref_domain_dispersion_store(h) bind(C, name='ref_domain_dispersion_store')
        integer(c_int), value :: h
        call pool(h)%copy_U_to_dispersive_last_timestep()
    end subroutine

    subroutine ref_domain_dispersion_append(h) bind(C, name='ref_domain_dispersion_append')
        integer(c_int), value :: h
        call pool(h)%ds%qet%append_value_and_time_for_quadratic_extrapolation(pool(h)%time, pool(h)%U(:,:,UH:VH))
    end subroutine

    subroutine ref_domain_dispersion_solve(h, rhs_up, estimate, ftime) bind(C, name='ref_domain_dispersion_solve')
        integer(c_int), value :: h, rhs_up, estimate
        real(c_double), value :: ftime
        call pool(h)%solve_dispersive_terms(rhs_is_up_to_date=(rhs_up /= 0), &
            estimate_solution_forward_in_time=(estimate /= 0), forward_time=ftime)
    end subroutine

    subroutine ref_domain_set_time(h, t) bind(C, name='ref_domain_set_time')
        integer(c_int), value :: h
        real(c_double), value :: t
        pool(h)%time = t
    end subroutine

    function ref_domain_volume_interior(h) result(v) bind(C, name='ref_domain_volume_interior')
        integer(c_int), value :: h
        real(c_double) :: v
        v = pool(h)%volume_interior()
    end function

    subroutine ref_domain_statistics(h, out) bind(C, name='ref_domain_statistics')
        ! compute_domain_statistics (domain_mod.f90:683-849): max/min stage, max/min speed,
        ! potential, kinetic and total energy / rho
        integer(c_int), value :: h
        real(c_double), intent(out) :: out(7)
        real(dp) :: v(7)
        call pool(h)%compute_domain_statistics(v(1), v(2), v(3), v(4), v(5), v(6), v(7))
        out = real(v, c_double)
    end subroutine

    function ref_domain_stationary_timestep_max(h) result(v) bind(C, name='ref_domain_stationary_timestep_max')
        integer(c_int), value :: h
        real(c_double) :: v
        v = pool(h)%stationary_timestep_max()
    end function

    subroutine ref_domain_get_max_U(h, k, out) bind(C, name='ref_domain_get_max_U')
        ! k (1-based) indexes [max_stage, max_speed, max_flux, arrival_time, time_of_max_stage, min_stage]
        integer(c_int), value :: h, k
        real(c_double), intent(out) :: out(*)
        integer :: n
        n = pool(h)%nx(1) * pool(h)%nx(2)
        out(1:n) = real(reshape(pool(h)%max_U(:,:,k), [n]), c_double)
    end subroutine

    function ref_stage_fun(domain, t, i, j) result(stage_uh_vh_elev)
        ! Sinusoidal stage boundary value; elevation is left as it is in the cell.
        type(domain_type), intent(in) :: domain
        real(dp), intent(in) :: t
        integer(ip), intent(in) :: i, j
        real(dp) :: stage_uh_vh_elev(4)
        integer :: h
        real(dp), parameter :: two_pi = 2.0_dp * acos(-1.0_dp)
        h = int(domain%local_index)
        stage_uh_vh_elev(1) = bc_msl(h) + bc_amp(h) * sin(two_pi * t / bc_period(h))
        stage_uh_vh_elev(2) = 0.0_dp
        stage_uh_vh_elev(3) = 0.0_dp
        stage_uh_vh_elev(4) = domain%U(i, j, ELV)
    end function

    subroutine ref_domain_set_stage_bc(h, amp, period, msl, sides) bind(C, name='ref_domain_set_stage_bc')
        ! Installs boundary_stage_transmissive_normal_momentum
        ! (propagation/SWALS/src/shallow_water/boundary_mod.f90:78-164) on the sides flagged in
        ! sides(1:4) = [N, E, S, W], with stage = msl + amp*sin(2*pi*t/period).
        integer(c_int), value :: h
        real(c_double), value :: amp, period, msl
        integer(c_int), intent(in) :: sides(4)
        bc_amp(h) = amp; bc_period(h) = period; bc_msl(h) = msl
        pool(h)%local_index = h
        pool(h)%boundary_exterior = (sides /= 0)
        pool(h)%boundary_function => ref_stage_fun
        pool(h)%boundary_subroutine => boundary_stage_transmissive_normal_momentum
    end subroutine

    subroutine ref_domain_set_flather_bc(h, use_fun, amp, period, msl, sides) bind(C, name='ref_domain_set_flather_bc')
        ! Installs flather_boundary (propagation/SWALS/src/shallow_water/boundary_mod.f90:502-720)
        ! on the sides flagged in sides(1:4) = [N, E, S, W]; with use_fun /= 0 the outside
        ! state comes from the sine stage function, else from the sea at rest.
        integer(c_int), value :: h, use_fun
        real(c_double), value :: amp, period, msl
        integer(c_int), intent(in) :: sides(4)
        bc_amp(h) = amp; bc_period(h) = period; bc_msl(h) = msl
        pool(h)%local_index = h
        pool(h)%boundary_exterior = (sides /= 0)
        if(use_fun /= 0) then
            pool(h)%boundary_function => ref_stage_fun
        else
            pool(h)%boundary_function => null()
        end if
        pool(h)%boundary_subroutine => flather_boundary
    end subroutine

    subroutine ref_domain_set_bc(h, kind, use_fun, amp, period, msl, sides) bind(C, name='ref_domain_set_bc')
        ! kind 3 transmissive_boundary, 4 boundary_stage_radiation_momentum, 5
        ! boundary_stage_transmissive_momentum (boundary_mod.f90:23-75, 166-405) on the sides
        ! flagged in sides(1:4) = [N, E, S, W]; use_fun /= 0: the sine stage function
        integer(c_int), value :: h, kind, use_fun
        real(c_double), value :: amp, period, msl
        integer(c_int), intent(in) :: sides(4)
        bc_amp(h) = amp; bc_period(h) = period; bc_msl(h) = msl
        pool(h)%local_index = h
        pool(h)%boundary_exterior = (sides /= 0)
        if(use_fun /= 0) then
            pool(h)%boundary_function => ref_stage_fun
        else
            pool(h)%boundary_function => null()
        end if
        select case(kind)
        case(3); pool(h)%boundary_subroutine => transmissive_boundary
        case(4); pool(h)%boundary_subroutine => boundary_stage_radiation_momentum
        case(5); pool(h)%boundary_subroutine => boundary_stage_transmissive_momentum
        end select
    end subroutine

    ! ---------------------------------------------------------------------
    ! Multidomain: setup through the reference's own md%setup, stepping with
    ! md%evolve_one_step (multidomain_mod.f90:1081-1212), and access to the
    ! nesting tables that setup builds (nested_grid_comms_mod.f90:69-219).
    ! ---------------------------------------------------------------------

    subroutine ref_md_begin(nd) bind(C, name='ref_md_begin')
        integer(c_int), value :: nd
        type(multidomain_type) :: fresh
        md = fresh
        allocate(md%domains(nd))
    end subroutine

    subroutine ref_md_set_domain(k, nx, ny, lw, ll, method, trf, dxrf) bind(C, name='ref_md_set_domain')
        integer(c_int), value :: k, nx, ny, method, trf
        real(c_double), intent(in) :: lw(2), ll(2)
        real(c_double), value :: dxrf
        md%domains(k)%lower_left = real(ll, dp)
        md%domains(k)%lw = real(lw, dp)
        md%domains(k)%nx = [nx, ny]
        md%domains(k)%dx = md%domains(k)%lw / md%domains(k)%nx
        md%domains(k)%timestepping_refinement_factor = trf
        md%domains(k)%dx_refinement_factor = real(dxrf, dp)
        md%domains(k)%timestepping_method = method_name(method)
    end subroutine

    subroutine ref_md_set_nested(k, parent, ll, ur, dxrf, trf, method) bind(C, name='ref_md_set_nested')
        integer(c_int), value :: k, parent, trf, method
        real(c_double), intent(in) :: ll(2), ur(2)
        real(c_double), value :: dxrf
        call md%domains(k)%match_geometry_to_parent(parent_domain=md%domains(parent), &
            lower_left=real(ll, dp), upper_right=real(ur, dp), dx_refinement_factor=int(dxrf, ip), &
            timestepping_refinement_factor=trf)
        md%domains(k)%timestepping_method = method_name(method)
    end subroutine

    subroutine ref_md_set_params(k, truely_linear, msl, lfc) bind(C, name='ref_md_set_params')
        integer(c_int), value :: k, truely_linear
        real(c_double), value :: msl, lfc
        md%domains(k)%linear_solver_is_truely_linear = (truely_linear /= 0)
        md%domains(k)%msl_linear = msl
        md%domains(k)%linear_friction_coeff = lfc
    end subroutine

    subroutine ref_md_setup(max_stage) bind(C, name='ref_md_setup')
        ! max_stage /= 0: every domain tracks max_stage (update_max_quantities)
        integer(c_int), value :: max_stage
        integer :: k
        do k = 1, size(md%domains)
            md%domains(k)%record_max_U = (max_stage /= 0)
            if(max_stage /= 0) then
                md%domains(k)%nontemporal_grids_to_store = [character(len=charlen):: 'max_stage']
            else
                md%domains(k)%nontemporal_grids_to_store = ['']
            end if
        end do
        call md%setup(verbose=.false.)
    end subroutine

    subroutine ref_md_sample(k, n, ij, out) bind(C, name='ref_md_sample')
        ! [STG, UH, VH] of domain k at n cells (1-based i, j pairs): out((var - 1) * n + g)
        integer(c_int), value :: k, n
        integer(c_int), intent(in) :: ij(2, *)
        real(c_double), intent(out) :: out(*)
        integer :: g, v
        do v = 1, 3
            do g = 1, n
                out((v - 1) * n + g) = real(md%domains(k)%U(ij(1, g), ij(2, g), v), c_double)
            end do
        end do
    end subroutine

    subroutine ref_md_get_max_stage(k, out) bind(C, name='ref_md_get_max_stage')
        integer(c_int), value :: k
        real(c_double), intent(out) :: out(*)
        integer :: n
        n = md%domains(k)%nx(1) * md%domains(k)%nx(2)
        out(1:n) = real(reshape(md%domains(k)%max_U(:, :, 1), [n]), c_double)
    end subroutine

    subroutine ref_md_dims(k, out) bind(C, name='ref_md_dims')
        ! out = nx, ny, n_send_comms, n_recv_comms, staggered, method index,
        !       timestepping_refinement_factor, exterior_cells_width
        integer(c_int), value :: k
        integer(c_int), intent(out) :: out(8)
        associate(d => md%domains(k))
            out(1:2) = d%nx
            out(3) = 0; out(4) = 0
            if(allocated(d%nesting%send_comms)) out(3) = size(d%nesting%send_comms)
            if(allocated(d%nesting%recv_comms)) out(4) = size(d%nesting%recv_comms)
            out(5) = merge(1, 0, d%is_staggered_grid)
            out(6) = timestepping_method_index(d%timestepping_method)
            out(7) = d%timestepping_refinement_factor
            out(8) = d%exterior_cells_width
        end associate
    end subroutine

    subroutine ref_md_geometry(k, lw, ll, dx) bind(C, name='ref_md_geometry')
        integer(c_int), value :: k
        real(c_double), intent(out) :: lw(2), ll(2), dx(2)
        lw = md%domains(k)%lw; ll = md%domains(k)%lower_left; dx = md%domains(k)%dx
    end subroutine

    subroutine ref_md_metrics(k, dbe, dle, area, cor, cor_be, tanlat) bind(C, name='ref_md_metrics')
        integer(c_int), value :: k
        real(c_double), intent(out) :: dbe(*), dle(*), area(*), cor(*), cor_be(*), tanlat(*)
        integer :: nx, ny
        associate(d => md%domains(k))
            nx = d%nx(1); ny = d%nx(2)
            dbe(1:(ny+1)) = d%distance_bottom_edge
            dle(1:(nx+1)) = d%distance_left_edge
            area(1:ny) = d%area_cell_y
            cor(1:ny) = 0.0_c_double; cor_be(1:(ny+1)) = 0.0_c_double; tanlat(1:ny) = 0.0_c_double
            if(allocated(d%coriolis)) cor(1:ny) = d%coriolis
            if(allocated(d%coriolis_bottom_edge)) cor_be(1:(ny+1)) = d%coriolis_bottom_edge
            if(allocated(d%tanlat_on_radius_earth)) tanlat(1:ny) = d%tanlat_on_radius_earth
        end associate
    end subroutine

    subroutine ref_md_set_U(k, U) bind(C, name='ref_md_set_U')
        integer(c_int), value :: k
        real(c_double), intent(in) :: U(*)
        integer :: n
        n = md%domains(k)%nx(1) * md%domains(k)%nx(2)
        md%domains(k)%U = reshape(real(U(1:4*n), dp), [md%domains(k)%nx(1), md%domains(k)%nx(2), 4])
    end subroutine

    subroutine ref_md_get_U(k, U) bind(C, name='ref_md_get_U')
        integer(c_int), value :: k
        real(c_double), intent(out) :: U(*)
        integer :: n
        n = md%domains(k)%nx(1) * md%domains(k)%nx(2)
        U(1:4*n) = real(reshape(md%domains(k)%U(:,:,1:4), [4*n]), c_double)
    end subroutine

    subroutine ref_md_set_manning_squared(k, m) bind(C, name='ref_md_set_manning_squared')
        integer(c_int), value :: k
        real(c_double), intent(in) :: m(*)
        integer :: n
        n = md%domains(k)%nx(1) * md%domains(k)%nx(2)
        md%domains(k)%manning_squared = reshape(real(m(1:n), dp), [md%domains(k)%nx(1), md%domains(k)%nx(2)])
    end subroutine

    subroutine ref_md_make_consistent() bind(C, name='ref_md_make_consistent')
        call md%make_initial_conditions_consistent()
    end subroutine

    subroutine ref_md_evolve(dt, nsteps) bind(C, name='ref_md_evolve')
        real(c_double), value :: dt
        integer(c_int), value :: nsteps
        integer :: n
        do n = 1, nsteps
            call md%evolve_one_step(real(dt, dp))
        end do
    end subroutine

    subroutine ref_md_statistics(k, out) bind(C, name='ref_md_statistics')
        ! domain k's compute_domain_statistics inside the multidomain (priority cells only)
        integer(c_int), value :: k
        real(c_double), intent(out) :: out(7)
        real(dp) :: v(7)
        call md%domains(k)%compute_domain_statistics(v(1), v(2), v(3), v(4), v(5), v(6), v(7))
        out = real(v, c_double)
    end subroutine

    subroutine ref_md_time(k, t) bind(C, name='ref_md_time')
        integer(c_int), value :: k
        real(c_double), intent(out) :: t
        t = md%domains(k)%time
    end subroutine

    subroutine ref_md_priority(k, pdi, pnp) bind(C, name='ref_md_priority')
        ! priority_domain_index(nx, ny) and is_priority_domain_not_periodic(nx, ny)
        integer(c_int), value :: k
        integer(c_int), intent(out) :: pdi(*), pnp(*)
        integer :: n
        n = md%domains(k)%nx(1) * md%domains(k)%nx(2)
        pdi(1:n) = reshape(md%domains(k)%nesting%priority_domain_index, [n])
        pnp(1:n) = reshape(md%domains(k)%nesting%is_priority_domain_not_periodic, [n])
    end subroutine

    subroutine ref_md_all(dx, methods) bind(C, name='ref_md_all')
        ! all_dx_md(1:2, k, 1) and all_timestepping_methods_md(k, 1), k = 1..nd
        real(c_double), intent(out) :: dx(*)
        integer(c_int), intent(out) :: methods(*)
        integer :: k, nd
        nd = size(md%domains)
        do k = 1, nd
            dx(2*k-1:2*k) = md%all_dx_md(1:2, k, 1)
            methods(k) = md%all_timestepping_methods_md(k, 1)
        end do
    end subroutine

    subroutine ref_md_comm(k, is_recv, c, iout, dout) bind(C, name='ref_md_comm')
        ! iout(1:6)   send_inds or recv_inds (iL, iU, jL, jU, kL, kU)
        ! iout(7:10)  my_domain_index, neighbour_domain_index, my_domain_comms_index,
        !             neighbour_domain_comms_index
        ! iout(11:15) my_domain_is_finer, equal_cell_ratios, my_domain_staggered_grid,
        !             neighbour_domain_staggered_grid, use_wetdry_limiting
        ! iout(16:18) buffer size, flux start, time index
        ! iout(19:26) (rows, nvar) of the box flux integral for N, S, E, W
        ! iout(27)    active
        ! dout(1:2)   cell_ratios
        integer(c_int), value :: k, is_recv, c
        integer(c_int), intent(out) :: iout(32)
        real(c_double), intent(out) :: dout(4)
        integer :: d
        iout = 0; dout = 0.0_c_double
        if(is_recv /= 0) then
            associate(cm => md%domains(k)%nesting%recv_comms(c))
                iout(1:6) = reshape(cm%recv_inds, [6])
                iout(16) = size(cm%recv_buffer); iout(17) = cm%recv_flux_start; iout(18) = cm%recv_timeind
                do d = 1, 4
                    iout(17 + 2*d) = size(cm%recv_box_flux_integral(d)%x, 1)
                    iout(18 + 2*d) = size(cm%recv_box_flux_integral(d)%x, 2)
                end do
                iout(27) = merge(1, 0, cm%recv_active)
                call common(cm)
            end associate
        else
            associate(cm => md%domains(k)%nesting%send_comms(c))
                iout(1:6) = reshape(cm%send_inds, [6])
                iout(16) = size(cm%send_buffer); iout(17) = cm%send_flux_start; iout(18) = cm%send_timeind
                do d = 1, 4
                    iout(17 + 2*d) = size(cm%send_box_flux_integral(d)%x, 1)
                    iout(18 + 2*d) = size(cm%send_box_flux_integral(d)%x, 2)
                end do
                iout(27) = merge(1, 0, cm%send_active)
                call common(cm)
            end associate
        end if
    contains
        subroutine common(cm)
            use nested_grid_comms_mod, only: two_way_nesting_comms_type
            type(two_way_nesting_comms_type), intent(in) :: cm
            iout(7) = cm%my_domain_index; iout(8) = cm%neighbour_domain_index
            iout(9) = cm%my_domain_comms_index; iout(10) = cm%neighbour_domain_comms_index
            iout(11) = merge(1, 0, cm%my_domain_is_finer); iout(12) = merge(1, 0, cm%equal_cell_ratios)
            iout(13) = cm%my_domain_staggered_grid; iout(14) = cm%neighbour_domain_staggered_grid
            iout(15) = merge(1, 0, cm%use_wetdry_limiting)
            dout(1:2) = cm%cell_ratios
        end subroutine
    end subroutine

    subroutine ref_md_comm_buffer(k, is_recv, c, out) bind(C, name='ref_md_comm_buffer')
        integer(c_int), value :: k, is_recv, c
        real(c_double), intent(out) :: out(*)
        if(is_recv /= 0) then
            out(1:size(md%domains(k)%nesting%recv_comms(c)%recv_buffer)) = md%domains(k)%nesting%recv_comms(c)%recv_buffer
        else
            out(1:size(md%domains(k)%nesting%send_comms(c)%send_buffer)) = md%domains(k)%nesting%send_comms(c)%send_buffer
        end if
    end subroutine

    subroutine ref_md_volume(out) bind(C, name='ref_md_volume')
        ! per-domain volume in the priority-domain cells (volume_in_priority_domain)
        real(c_double), intent(out) :: out(*)
        integer :: k
        real(force_double) :: v
        do k = 1, size(md%domains)
            call md%domains(k)%volume_in_priority_domain(v)
            out(k) = v
        end do
    end subroutine

end module
