!---------------------------------------------------------------------------
! simp_binding.f90 -- iso_c_binding interfaces of include/simp.h.
!
! A Fortran SIMP build uses this module to reach the MI355X library
! (libsimp_amd.so).  Arrays follow the reference's module layout
! (modules/data_variables.f90): x(dim, n) interleaved, REAL(8), and are
! passed as assumed-size actual arguments.
!---------------------------------------------------------------------------
module simp_binding
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: SIMP_OK = 0
    integer(c_int), parameter :: SIMP_FP64 = 0, SIMP_MIXED = 1

    type, bind(C) :: simp_config
        integer(c_int32_t) :: version, precision, int_algorithm,          &
                              nnps_algorithm, sle, tse, rheology, bfa,     &
                              eos_alg, pa_sph, skf, den_alg, wind_mode,    &
                              max_particles, max_boundary, reserved
        real(c_double)     :: wind(2), current(2), max_hsml
        integer(c_int32_t) :: artificial_visc, art_visc_alg,               &
                              xsph_regularization, boundary_friction,      &
                              boundary_pressure, pressure_force,           &
                              wind_force, water_force, viscosity
    end type simp_config

    interface
        subroutine simp_config_default(cfg) bind(C, name='simp_config_default')
            import :: simp_config
            type(simp_config), intent(out) :: cfg
        end subroutine simp_config_default

        integer(c_int) function simp_create(cfg, h) bind(C, name='simp_create')
            import :: simp_config, c_ptr, c_int
            type(simp_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: h
        end function simp_create

        integer(c_int) function simp_destroy(h) bind(C, name='simp_destroy')
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function simp_destroy

        integer(c_int) function simp_single_step(h, n, x, vx, mass, rho,     &
                hsml, hth, A, nb, xb, massb, current_time, itimestep, dxdt,  &
                dvxdt, dhdt, dAdt, eps3, zeta, eta, p, sig3, nip, bnip)      &
                bind(C, name='simp_single_step')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: n, nb, itimestep
            real(c_double), intent(in) :: x(*), vx(*), mass(*), rho(*),     &
                                          hsml(*), hth(*), A(*), xb(*),     &
                                          massb(*)
            real(c_double), value :: current_time
            real(c_double), intent(out) :: dxdt(*), dvxdt(*), dhdt(*),     &
                                           dAdt(*), eps3(*), zeta(*),      &
                                           eta(*), p(*), sig3(*)
            integer(c_int), intent(out) :: nip, bnip
        end function simp_single_step

        integer(c_int) function simp_kernel_values(h, cap, nip, p_i, p_j,    &
                w, kernel_zero) bind(C, name='simp_kernel_values')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: cap
            integer(c_int), intent(out) :: nip, p_i(*), p_j(*)
            real(c_double), intent(out) :: w(*), kernel_zero(*)
        end function simp_kernel_values

        integer(c_int) function simp_set_module_state(h, dt, nb, b_normal) &
                bind(C, name='simp_set_module_state')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            real(c_double), value :: dt
            integer(c_int), value :: nb
            real(c_double), intent(in) :: b_normal(*)
        end function simp_set_module_state

        integer(c_int) function simp_kernel_gradients(h, cap, nip, p_i, p_j, &
                dwdx) bind(C, name='simp_kernel_gradients')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: cap
            integer(c_int), intent(out) :: nip, p_i(*), p_j(*)
            real(c_double), intent(out) :: dwdx(*)
        end function simp_kernel_gradients

        integer(c_int) function simp_upload(h, n, x, vx, mass, rho, hsml,   &
                hth, A, nb, xb, massb, dt, current_time, itimestep)          &
                bind(C, name='simp_upload')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: n, nb, itimestep
            real(c_double), intent(in) :: x(*), vx(*), mass(*), rho(*),     &
                                          hsml(*), hth(*), A(*), xb(*),     &
                                          massb(*)
            real(c_double), value :: dt, current_time
        end function simp_upload

        integer(c_int) function simp_step_pc(h, nsteps, sync)               &
                bind(C, name='simp_step_pc')
            import :: c_ptr, c_int
            type(c_ptr), value :: h
            integer(c_int), value :: nsteps, sync
        end function simp_step_pc

        integer(c_int) function simp_download(h, x, vx, rho, hsml, hth, A,  &
                dt, current_time, itimestep) bind(C, name='simp_download')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: x(*), vx(*), rho(*), hsml(*),    &
                                           hth(*), A(*)
            real(c_double), intent(out) :: dt, current_time
            integer(c_int), intent(out) :: itimestep
        end function simp_download

        ! snapshots in the reference's CSV format (output.f90) and exact
        ! binary snapshots; folder: NUL-terminated (trim(path)//c_null_char)
        integer(c_int) function simp_output_create(folder)                  &
                bind(C, name='simp_output_create')
            import :: c_int, c_char
            character(kind=c_char), intent(in) :: folder(*)
        end function simp_output_create

        integer(c_int) function simp_output_append(folder, n, x, vx, mass,  &
                p, hsml, A, hth, eps3, sig3, nb, current_time)              &
                bind(C, name='simp_output_append')
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: folder(*)
            integer(c_int), value :: n, nb
            real(c_double), intent(in) :: x(*), vx(*), mass(*), p(*),       &
                                          hsml(*), A(*), hth(*), eps3(*),   &
                                          sig3(*)
            real(c_double), value :: current_time
        end function simp_output_append

        integer(c_long_long) function simp_snapshot_size(h)                &
                bind(C, name='simp_snapshot_size')
            import :: c_ptr, c_long_long
            type(c_ptr), value :: h
        end function simp_snapshot_size

        integer(c_int) function simp_snapshot_save(h, buf, cap)             &
                bind(C, name='simp_snapshot_save')
            import :: c_ptr, c_int, c_long_long
            type(c_ptr), value :: h, buf
            integer(c_long_long), value :: cap
        end function simp_snapshot_save

        integer(c_int) function simp_snapshot_load(h, buf, len)             &
                bind(C, name='simp_snapshot_load')
            import :: c_ptr, c_int, c_long_long
            type(c_ptr), value :: h, buf
            integer(c_long_long), value :: len
        end function simp_snapshot_load

        type(c_ptr) function simp_strerror(status) bind(C, name='simp_strerror')
            import :: c_ptr, c_int
            integer(c_int), value :: status
        end function simp_strerror
    end interface

contains

    ! WRITE + STOP on a nonzero status, as the reference reports capacity
    ! overflow (neighbor/nnps.f90:415-423).
    subroutine simp_check(status, where)
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: i
        character(len=64) :: text
        if (status == SIMP_OK) return
        call c_f_pointer(simp_strerror(status), msg, [64])
        text = ' '
        do i = 1, 64
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        enddo
        write(*,*) '******************************************************'
        write(*,*) ' ', where, ': ', trim(text)
        write(*,*) '******************************************************'
        stop
    end subroutine simp_check

end module simp_binding
