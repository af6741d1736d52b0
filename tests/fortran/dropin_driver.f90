! Drop-in check through the Fortran boundary: the reference's own domain_type
! (compiled from the reference sources into oracle/_ref/) is set up as a SWALS
! application would (allocate_quantities, initial conditions), then the same
! domain is advanced on the GPU through swals_hip_mod (c_loc of the domain's own
! metric arrays and U) and on the CPU with domain%evolve_one_step. Prints the max
! |difference| of each U plane and 'BITWISE_EQUAL' when every plane matches exactly.
!
! Usage: dropin_driver <method> <nsteps>  (method: linear | llpnf | nonlinear)
!
! 'nonlinear' runs 'leapfrog_nonlinear' without walls, with the reference's Flather
! boundary (domain%boundary_subroutine => flather_boundary) and a co-seismic forcing
! patch (forcing_patch_type%setup + apply_forcing_patch), each mapped to the backend
! (swh_set_flather_boundary; swh_set_forcing_patch with the patch's own forcing_work).
program dropin_driver
    use iso_c_binding
    use global_mod, only: dp, ip, charlen
    use domain_mod, only: domain_type, STG, UH, VH, ELV
    use boundary_mod, only: flather_boundary
    use forcing_mod, only: forcing_patch_type, apply_forcing_patch
    use swals_hip_mod
    implicit none

    type(domain_type), target :: dom
    type(forcing_patch_type), target :: fp
    type(swh_domain_desc) :: desc
    type(c_ptr) :: h
    real(dp), allocatable, target :: Ugpu(:,:,:)
    real(dp) :: dt, lw(2), ll(2), r2, diff(4)
    integer(ip) :: nx, ny, i, j, k, nsteps
    integer(c_int32_t) :: sides(4)
    character(len=64) :: arg
    logical :: llpnf, nonlinear

    call get_command_argument(1, arg)
    llpnf = (trim(arg) == 'llpnf')
    nonlinear = (trim(arg) == 'nonlinear')
    call get_command_argument(2, arg)
    read(arg, *) nsteps

    nx = 96; ny = 80
    ll = [140.0_dp, -40.0_dp]
    lw = [16.0_dp, 13.333333333333334_dp]
    if(llpnf) then
        dom%timestepping_method = 'leapfrog_linear_plus_nonlinear_friction'
    else if(nonlinear) then
        dom%timestepping_method = 'leapfrog_nonlinear'
    else
        dom%timestepping_method = 'linear'
    end if
    dom%nontemporal_grids_to_store = [character(len=charlen):: 'max_stage']
    call dom%allocate_quantities(lw, [nx, ny], ll, create_output_files=.false., verbose=.false.)

    ! bathymetry: a sloping sea floor rising to land in the east, walls all round
    do j = 1, ny
        do i = 1, nx
            dom%U(i, j, ELV) = -4000.0_dp + 4100.0_dp * real(i, dp) / real(nx, dp) &
                + 300.0_dp * sin(0.3_dp * j) * cos(0.2_dp * i)
            r2 = ((dom%x(i) - 146.0_dp)**2 + (dom%y(j) + 33.0_dp)**2) / 0.8_dp**2
            dom%U(i, j, STG) = merge(exp(-r2), 0.0_dp, dom%U(i, j, ELV) < 0.0_dp)
        end do
    end do
    if(.not. nonlinear) then
        dom%U(1, :, ELV) = 1.0e6_dp; dom%U(nx, :, ELV) = 1.0e6_dp
        dom%U(:, 1, ELV) = 1.0e6_dp; dom%U(:, ny, ELV) = 1.0e6_dp
    end if
    dom%U(:, :, UH:VH) = 0.0_dp
    if(llpnf) dom%manning_squared = 0.03_dp**2
    dt = dom%stationary_timestep_max()
    if(nonlinear) then
        ! open boundaries and an uplift over 10 steps starting inside the first
        dom%boundary_subroutine => flather_boundary
        call fp%setup(0.4_dp * dt, 10.4_dp * dt, 30_ip, 50_ip, 20_ip, 45_ip, STG, ELV)
        do j = 20, 45
            do i = 30, 50
                r2 = ((i - 40.0_dp)**2 + (j - 32.0_dp)**2) / 40.0_dp
                fp%forcing_work(i, j, STG) = 2.0_dp * exp(-r2)
                fp%forcing_work(i, j, UH) = 0.0_dp
                fp%forcing_work(i, j, VH) = 0.0_dp
                fp%forcing_work(i, j, ELV) = 2.0_dp * exp(-r2)
            end do
        end do
        dom%forcing_context_cptr = c_loc(fp)
        dom%forcing_subroutine => apply_forcing_patch
    end if

    ! --- GPU: the same domain handed over through the C ABI ---
    desc%nx = nx; desc%ny = ny
    desc%precision = SWH_PREC_FP64
    if(llpnf) then
        desc%timestepping_method = SWH_METHOD_LLPNF
    else if(nonlinear) then
        desc%timestepping_method = SWH_METHOD_LEAPFROG_NONLINEAR
    else
        desc%timestepping_method = SWH_METHOD_LINEAR
    end if
    desc%linear_solver_is_truely_linear = merge(1, 0, dom%linear_solver_is_truely_linear)
    desc%include_coriolis = 1
    desc%spherical = 1
    desc%exterior_cells_width = dom%exterior_cells_width
    desc%max_U_mask = SWH_MAXQ_MAX_STAGE
    desc%msl_linear = dom%msl_linear
    desc%cfl = dom%cfl
    desc%linear_friction_coeff = dom%linear_friction_coeff
    desc%distance_bottom_edge = c_loc(dom%distance_bottom_edge)
    desc%distance_left_edge = c_loc(dom%distance_left_edge)
    desc%area_cell_y = c_loc(dom%area_cell_y)
    desc%coriolis = c_loc(dom%coriolis)
    desc%coriolis_bottom_edge = c_loc(dom%coriolis_bottom_edge)
    desc%tanlat_on_radius_earth = c_loc(dom%tanlat_on_radius_earth)
    call swh_check(swh_domain_create(desc, h), 'swh_domain_create')
    if(llpnf .or. nonlinear) &
        call swh_check(swh_upload_manning_squared(h, dom%manning_squared), 'swh_upload_manning_squared')
    call swh_check(swh_upload_U(h, dom%U), 'swh_upload_U')
    if(nonlinear) then
        sides = merge(1_c_int32_t, 0_c_int32_t, dom%boundary_exterior)
        call swh_check(swh_set_flather_boundary(h, sides, 0_c_int32_t, 0.0_dp, 1.0_dp, 0.0_dp), 'flather')
        call swh_check(swh_set_forcing_patch(h, fp%start_time, fp%end_time, int(fp%i0, c_int32_t), &
            int(fp%i1, c_int32_t), int(fp%j0, c_int32_t), int(fp%j1, c_int32_t), int(fp%k0, c_int32_t), &
            int(fp%k1, c_int32_t), fp%forcing_work), 'forcing')
    end if
    call swh_check(swh_evolve_steps(h, dt, int(nsteps, c_int32_t)), 'swh_evolve_steps')
    allocate(Ugpu(nx, ny, 4))
    call swh_check(swh_download_U(h, Ugpu), 'swh_download_U')
    call swh_check(swh_domain_destroy(h), 'swh_domain_destroy')

    ! --- CPU: the reference's own evolve loop ---
    do k = 1, nsteps
        call dom%evolve_one_step(dt)
    end do

    do k = 1, 4
        diff(k) = maxval(abs(Ugpu(:, :, k) - dom%U(:, :, k)))
    end do
    write(*, '(A, 4ES12.4)') 'max |gpu - reference| per plane: ', diff
    if(all(Ugpu == dom%U)) then
        write(*, *) 'BITWISE_EQUAL'
    else
        write(*, *) 'DIFFERENT'
    end if
end program
