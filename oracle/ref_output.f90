!---------------------------------------------------------------------------
! ref_output.f90 -- driver that calls the reference's own output routine
! (in_out/output.f90) on states read from a binary file, to write golden
! CSV snapshots.
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference's objects (compiled
! in place by oracle/build_ref.sh into oracle/_ref/).  The files are
! created empty first, as initialize.f90:83-94 does; output appends to
! $PWD // output_folder (globals.f90:203-208).
!
! Input (stream, little endian):
!   int32  n, nb, snapshots
!   per snapshot:
!     real8  current_time
!     real8  x(2,n), vx(2,n), mass(n), p(n), hsml(n), A(n), h(n),
!            epsilon_11(n), epsilon_12(n), epsilon_22(n),
!            sigma_11(n), sigma_12(n), sigma_22(n)
! Arguments: input file, output folder relative to $PWD (with its leading
! and trailing '/').
!---------------------------------------------------------------------------
program simp_ref_output
    use globals
    use data_variables
    implicit none

    character(len=4096) :: fin, folder
    integer :: n, nb, snaps, s, u

    call get_command_argument(1, fin)
    call get_command_argument(2, folder)
    call sys_variables
    output_folder = trim(folder)

    open(newunit=u, file=trim(fin), access='stream', form='unformatted', &
         status='old', action='read')
    read(u) n, nb, snaps
    if (n > maxn .or. n < 1) then
        write(*,*) 'ref_output: particle count outside the reference capacity'
        stop 2
    endif

    open(1, file=cwd // trim(output_folder) // "particle_momentum.csv", &
         action='write', status='replace')
    open(2, file=cwd // trim(output_folder) // "particle_state.csv", &
         action='write', status='replace')
    open(3, file=cwd // trim(output_folder) // "particle_other.csv", &
         action='write', status='replace')
    close(1)
    close(2)
    close(3)

    ntotal = n
    ntotalb = nb
    do s = 1, snaps
        read(u) current_time
        read(u) x(:, 1:n), vx(:, 1:n), mass(1:n), p(1:n), hsml(1:n), &
                A(1:n), h(1:n), epsilon_11(1:n), epsilon_12(1:n), &
                epsilon_22(1:n), sigma_11(1:n), sigma_12(1:n), sigma_22(1:n)
        call output
    enddo
    close(u)
end program simp_ref_output
