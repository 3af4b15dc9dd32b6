!---------------------------------------------------------------------------
! single_step_gpu.f90 -- replacement of SIMP's single_step
! (integration/single_step.f90:5-108) that runs the right-hand side on the
! MI355X through the C ABI.  Same name, no dummy arguments, same module
! state read and written, so time_integration (and every other caller)
! links against it unchanged.
!
! Reads  ntotal, x, vx, mass, rho, hsml, h, A, ntotalb, xb, massb,
!        current_time, itimestep, max_hsml; with bfa = 2 also dt and
!        b_normal (the Wang boundary force)
! Writes dxdt, dvxdt, dhdt, dAdt, epsilon_11/12/22, zeta, eta, p,
!        sigma_11/12/22, nip, bnip; with den_alg = 1 also p_i, p_j, w and
!        kernel_zero, which sum_density (density.f90:41-100) reads next;
!        with den_alg = 2, sle = 1 or the Monaghan time step with artificial
!        viscosity p_i, p_j and dwdx, which con_density (density.f90:107-165),
!        Benz_hsml (smoothing_length.f90:75-99) and Monaghan_time_step
!        (time_step.f90:89-107) read
! The other pair arrays (r_ij, hsml_ij, boundary pairs) stay on the
! device; no active caller reads them on the host.
!---------------------------------------------------------------------------
subroutine single_step
    use globals
    use data_variables
    use simp_binding
    implicit none

    type(c_ptr), save :: handle = c_null_ptr
    double precision, allocatable, save :: eps3(:, :), sig3(:, :)
    type(simp_config) :: cfg
    integer(c_int) :: st, nipc, bnipc

    if (.not. c_associated(handle)) then
        call simp_config_default(cfg)
        cfg%max_hsml = max_hsml
        cfg%den_alg = den_alg
        cfg%pa_sph = pa_sph
        cfg%eos_alg = eos_alg
        cfg%rheology = rheology
        cfg%skf = skf
        cfg%int_algorithm = int_algorithm
        cfg%tse = tse
        cfg%sle = sle
        cfg%bfa = bfa
        cfg%art_visc_alg = art_visc_alg
        cfg%artificial_visc = merge(1, 0, artificial_visc)
        cfg%xsph_regularization = merge(1, 0, xsph_regularization)
        cfg%boundary_friction = merge(1, 0, boundary_friction)
        cfg%boundary_pressure = merge(1, 0, boundary_pressure)
        cfg%pressure_force = merge(1, 0, pressure_force)
        cfg%wind_force = merge(1, 0, wind_force)
        cfg%water_force = merge(1, 0, water_force)
        ! without viscosity zeta and eta stay at initialize's zero, as the
        ! device path returns them (internal_stress.f90:20-31)
        cfg%viscosity = merge(1, 0, viscosity)
        st = simp_create(cfg, handle)
        call simp_check(st, 'simp_create')
    endif
    if (allocated(eps3)) then
        if (size(eps3, 2) /= ntotal) deallocate(eps3, sig3)
    endif
    if (.not. allocated(eps3)) allocate(eps3(3, ntotal), sig3(3, ntotal))

    if (bfa == 2) then
        st = simp_set_module_state(handle, dt, ntotalb, b_normal)
        call simp_check(st, 'single_step module state')
    endif
    st = simp_single_step(handle, ntotal, x, vx, mass, rho, hsml, h, A,    &
                          ntotalb, xb, massb, current_time, itimestep,     &
                          dxdt, dvxdt, dhdt, dAdt, eps3, zeta, eta, p,     &
                          sig3, nipc, bnipc)
    call simp_check(st, 'single_step')

    epsilon_11(:ntotal) = eps3(1, :)
    epsilon_12(:ntotal) = eps3(2, :)
    epsilon_22(:ntotal) = eps3(3, :)
    sigma_11(:ntotal) = sig3(1, :)
    sigma_12(:ntotal) = sig3(2, :)
    sigma_22(:ntotal) = sig3(3, :)
    nip = nipc
    bnip = bnipc
    if (den_alg == 1) then
        st = simp_kernel_values(handle, size(w), nipc, p_i, p_j, w,       &
                                kernel_zero)
        call simp_check(st, 'single_step kernel values')
    endif
    if (den_alg == 2 .or. sle == 1 .or. (tse == 1 .and. artificial_visc)) then
        ! con_density, Benz_hsml and the viscous time step read the pairs
        ! and their gradients
        st = simp_kernel_gradients(handle, size(p_i), nipc, p_i, p_j, dwdx)
        call simp_check(st, 'single_step kernel gradients')
    endif
end subroutine single_step
