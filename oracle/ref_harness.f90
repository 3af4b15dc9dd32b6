!---------------------------------------------------------------------------
! ref_harness.f90 -- driver that runs the reference SIMP routines on a
! state read from a binary file and writes their results back in FP64.
!
! TEST INFRASTRUCTURE ONLY.  It is linked with the reference's own objects
! (compiled in place by oracle/build_ref.sh into oracle/_ref/) and replaces
! simp.f90 (program SIMP) as the caller: it sets the module state the
! reference's initialize would set (initialize.f90:190-209), then calls the
! reference's external NNPS, single_step or time_integration unchanged.
!
! Input (stream, little endian):
!   int32  mode            1 = NNPS, 2 = single_step, 3 = time_integration,
!                          4 = single_step on a first state, then on a second
!                          state that follows the first in the file
!   int32  n, nb, reps     reps: number of consecutive single_step calls
!   real8  dt, current_time, integration_time, max_hsml
!   real8  x(2,n), vx(2,n), mass(n), rho(n), hsml(n), A(n), h(n)
!   real8  xb(2,nb), massb(nb), b_normal(2,nb)
! The sound speed is set from the densities read, as initialize does after
! setting them (initialize.f90:199-200); sound_speed computes it only for
! the switches that read it (sound_speed.f90:21).
! Output (stream):
!   int32  nip, bnip, itimestep, pad
!   real8  dt, current_time, wall_seconds
!   int32  p_i(nip), p_j(nip); real8 r_ij(nip), hsml_ij(nip)
!   int32  bp_i(bnip), bp_j(bnip); real8 br_ij(bnip)
!   real8  x(2,n), vx(2,n), rho(n), hsml(n), A(n), h(n)
!   real8  dxdt(2,n), dvxdt(2,n), dhdt(n), dAdt(n)
!   real8  epsilon_11(n), epsilon_12(n), epsilon_22(n), zeta(n), eta(n),
!          p(n), sigma_11(n), sigma_12(n), sigma_22(n)
!   real8  w(nip), kernel_zero(n)   (kernel values of the last single_step)
!---------------------------------------------------------------------------
program simp_ref_harness
    use globals
    use data_variables
    implicit none

    character(len=4096) :: fin, fout
    integer :: mode, n, nb, reps, pad, k, u
    integer(kind=8) :: c0, c1, crate
    double precision :: wall

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)

    open(newunit=u, file=trim(fin), access='stream', form='unformatted', &
         status='old', action='read')
    read(u) mode, n, nb, reps
    if (n > maxn .or. nb > maxnb .or. n < 1) then
        write(*,*) 'harness: particle count outside the reference capacity'
        stop 2
    endif
    read(u) dt, current_time, integration_time, max_hsml
    read(u) x(:, 1:n), vx(:, 1:n), mass(1:n), rho(1:n), hsml(1:n), &
            A(1:n), h(1:n)
    if (nb > 0) read(u) xb(:, 1:nb), massb(1:nb), b_normal(:, 1:nb)
    close(u)

    ntotal = n
    ntotalb = nb
    ! the kernel's scale factor as initialize sets it (initialize.f90:207-217):
    ! 1 for the Wendland kernels, 2 for the quadratic, 3 for the others
    select case (skf)
    case (4, 5, 6)
        scale_h = 1
    case (1)
        scale_h = 2
    case default
        scale_h = 3
    end select
    itimestep = 1
    call sound_speed

    ! the CSV files time_integration's output routine appends to
    output_folder = '/'
    call sys_variables
    open(newunit=u, file=cwd // '/particle_momentum.csv', status='replace')
    close(u)
    open(newunit=u, file=cwd // '/particle_state.csv', status='replace')
    close(u)
    open(newunit=u, file=cwd // '/particle_other.csv', status='replace')
    close(u)

    call system_clock(c0, crate)
    if (mode == 1) then
        call NNPS
    else if (mode == 2) then
        do k = 1, max(reps, 1)
            call single_step
        enddo
    else if (mode == 3) then
        call time_integration
    else if (mode == 4) then
        call single_step
        open(newunit=u, file=trim(fin), access='stream', form='unformatted', &
             status='old', action='read')
        read(u) k, k, k, k
        read(u) dt, current_time, integration_time, max_hsml
        read(u) x(:, 1:n), vx(:, 1:n), mass(1:n), rho(1:n), hsml(1:n), &
                A(1:n), h(1:n)
        if (nb > 0) read(u) xb(:, 1:nb), massb(1:nb), b_normal(:, 1:nb)
        read(u) x(:, 1:n), vx(:, 1:n), mass(1:n), rho(1:n), hsml(1:n), &
                A(1:n), h(1:n)
        if (nb > 0) read(u) xb(:, 1:nb), massb(1:nb), b_normal(:, 1:nb)
        close(u)
        call sound_speed
        call single_step
    else
        write(*,*) 'harness: unknown mode'
        stop 2
    endif
    call system_clock(c1)
    wall = dble(c1 - c0) / dble(crate)

    pad = 0
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', &
         status='replace', action='write')
    write(u) nip, bnip, itimestep, pad
    write(u) dt, current_time, wall
    write(u) p_i(1:nip), p_j(1:nip), r_ij(1:nip), hsml_ij(1:nip)
    write(u) bp_i(1:bnip), bp_j(1:bnip), br_ij(1:bnip)
    write(u) x(:, 1:n), vx(:, 1:n), rho(1:n), hsml(1:n), A(1:n), h(1:n)
    write(u) dxdt(:, 1:n), dvxdt(:, 1:n), dhdt(1:n), dAdt(1:n)
    write(u) epsilon_11(1:n), epsilon_12(1:n), epsilon_22(1:n), zeta(1:n), &
             eta(1:n), p(1:n), sigma_11(1:n), sigma_12(1:n), sigma_22(1:n)
    write(u) w(1:nip), kernel_zero(1:n)
    close(u)
end program simp_ref_harness
