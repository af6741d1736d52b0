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
    use boundary_mod, only: boundary_stage_transmissive_normal_momentum
    implicit none

    integer, parameter :: max_domains = 32
    type(domain_type), target, save :: pool(max_domains)
    logical, save :: in_use(max_domains) = .false.

    ! Parameters of the sinusoidal stage boundary (one set per handle)
    real(dp), save :: bc_amp(max_domains) = 0.0_dp, bc_period(max_domains) = 1.0_dp, &
        bc_msl(max_domains) = 0.0_dp

contains

    function method_name(method) result(name)
        ! Method ids follow the order of the reference's timestepping_metadata table
        ! (propagation/SWALS/src/shallow_water/timestepping_metadata_mod.f90:69-140).
        integer(c_int), intent(in) :: method
        character(len=charlen) :: name
        select case(method)
        case(1); name = 'rk2'
        case(4); name = 'euler'
        case(5); name = 'linear'
        case(6); name = 'leapfrog_linear_plus_nonlinear_friction'
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

end module
