!---------------------------------------------------------------------------
! ref_init.f90 -- driver that runs the reference's own initialize
! (in_out/initialize.f90: particle_placing from the domain map,
! boundaries -> boundary_placing_1 from the boundary map, time_step) and
! writes the maps, their geometry and the initial state it produced, for
! the golden vectors of the scenario generator.
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference's objects (compiled
! in place by oracle/build_ref.sh into oracle/_ref/).  initialize writes its
! empty CSV files into $PWD // output_folder, which must exist.
!
! Output (stream, little endian):
!   int32  ntotal, ntotalb, x_numcell, y_numcell, bx_numcell, by_numcell
!   int32  map(x_numcell, y_numcell), bmap(bx_numcell, by_numcell)
!   real8  x_mingeom, y_maxgeom, map_resolution, bx_mingeom, by_maxgeom,
!          bmap_resolution, map_factor, resolution, b_length,
!          hsml_multiplicity, rho_ice
!   real8  dt, max_hsml, current_time
!   real8  x(2,n), vx(2,n), mass(n), rho(n), hsml(n), A(n), h(n)
!   real8  xb(2,nb), massb(nb)
! Arguments: output file, output folder relative to $PWD (leading and
! trailing '/').
!---------------------------------------------------------------------------
program simp_ref_init
    use globals
    use data_variables
    use domain
    use boundary_domain
    implicit none

    character(len=4096) :: fout, folder
    integer :: u

    call get_command_argument(1, fout)
    call get_command_argument(2, folder)
    call sys_variables
    output_folder = trim(folder)
    call initialize

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', &
         status='replace', action='write')
    write(u) ntotal, ntotalb, x_numcell, y_numcell, bx_numcell, by_numcell
    write(u) map, bmap
    write(u) x_mingeom, y_maxgeom, map_resolution, bx_mingeom, by_maxgeom, &
             bmap_resolution, dble(map_factor), resolution, dble(b_length), &
             dble(hsml_multiplicity), dble(rho_ice)
    write(u) dt, max_hsml, current_time
    write(u) x(:, 1:ntotal), vx(:, 1:ntotal), mass(1:ntotal), rho(1:ntotal), &
             hsml(1:ntotal), A(1:ntotal), h(1:ntotal)
    if (ntotalb > 0) write(u) xb(:, 1:ntotalb), massb(1:ntotalb)
    close(u)
end program simp_ref_init
