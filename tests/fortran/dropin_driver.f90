! Drop-in check through the Fortran boundary: the reference's own domain_type
! (compiled from the reference sources into oracle/_ref/) is set up as a SWALS
! application would (allocate_quantities, initial conditions), then the same
! domain is advanced on the GPU through swals_hip_mod (c_loc of the domain's own
! metric arrays and U) and on the CPU with domain%evolve_one_step. Prints the max
! |difference| of each U plane and 'BITWISE_EQUAL' when every plane matches exactly.
!
! Usage: dropin_driver <method> <nsteps>  (method: linear | llpnf)
program dropin_driver
    use iso_c_binding
    use global_mod, only: dp, ip, charlen
    use domain_mod, only: domain_type, STG, UH, VH, ELV
    use swals_hip_mod
    implicit none

    type(domain_type), target :: dom
    type(swh_domain_desc) :: desc
    type(c_ptr) :: h
    real(dp), allocatable, target :: Ugpu(:,:,:)
    real(dp) :: dt, lw(2), ll(2), r2, diff(4)
    integer(ip) :: nx, ny, i, j, k, nsteps
    character(len=64) :: arg
    logical :: llpnf

    call get_command_argument(1, arg)
    llpnf = (trim(arg) == 'llpnf')
    call get_command_argument(2, arg)
    read(arg, *) nsteps

    nx = 96; ny = 80
    ll = [140.0_dp, -40.0_dp]
    lw = [16.0_dp, 13.333333333333334_dp]
    if(llpnf) then
        dom%timestepping_method = 'leapfrog_linear_plus_nonlinear_friction'
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
    dom%U(1, :, ELV) = 1.0e6_dp; dom%U(nx, :, ELV) = 1.0e6_dp
    dom%U(:, 1, ELV) = 1.0e6_dp; dom%U(:, ny, ELV) = 1.0e6_dp
    dom%U(:, :, UH:VH) = 0.0_dp
    if(llpnf) dom%manning_squared = 0.03_dp**2
    dt = dom%stationary_timestep_max()

    ! --- GPU: the same domain handed over through the C ABI ---
    desc%nx = nx; desc%ny = ny
    desc%precision = SWH_PREC_FP64
    if(llpnf) then
        desc%timestepping_method = SWH_METHOD_LLPNF
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
    if(llpnf) call swh_check(swh_upload_manning_squared(h, dom%manning_squared), 'swh_upload_manning_squared')
    call swh_check(swh_upload_U(h, dom%U), 'swh_upload_U')
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
