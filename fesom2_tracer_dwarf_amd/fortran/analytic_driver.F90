!> Fortran host driving the GPU library through the ISO_C_BINDING layer: the
!! analytic-dwarf workflow (src/fesom_analytic.F90) at the C-ABI level.
!!
!!   fesom_tracer_amd_driver nx ny nl nsteps
!!
!! Doubly periodic nx x ny mesh (Lx = Ly = 5 km per node), nl interfaces of
!! 10 m, bottom levels U{3..nl} (seed 1), T = 15 + 5 sin(2 pi x/Lx)
!! cos(2 pi y/Ly) - 10 z_n and S = 34 + (0.5 + 0.5 sin(2 pi x/Lx)) (1 - z_n),
!! MUSCL + QR4C + FCT, analytic velocity at CFL 0.3, nsteps steps of both
!! tracers; prints "step k: T min max sum" lines and the final S statistics.
program analytic_driver
  use iso_c_binding
  use fesom_tracer_amd
  implicit none
  integer :: nx, ny, nl, nsteps, n, k, st, nz, nn
  character(len=32) :: arg
  real(c_double) :: dt, tmin, tmax, tsum, d(7), lx, ly, x, y, zn
  real(c_double), allocatable :: coord(:), t(:, :), s(:, :)
  real(c_double), parameter :: pi = 3.14159265358979323846d0

  nx = 32; ny = 24; nl = 16; nsteps = 5
  if (command_argument_count() >= 4) then
    call get_command_argument(1, arg); read(arg, *) nx
    call get_command_argument(2, arg); read(arg, *) ny
    call get_command_argument(3, arg); read(arg, *) nl
    call get_command_argument(4, arg); read(arg, *) nsteps
  end if
  lx = 5.0d3 * nx
  ly = 5.0d3 * ny
  st = tracer_init_mpi(0_c_int)
  st = tracer_generate_mesh(int(nx, c_int), int(ny, c_int), 1_c_int, lx, ly, int(nl, c_int), 10.0d0, 0.0d0, &
                            3_c_int, int(nl, c_int), 1_c_long_long, 1_c_int)
  if (st /= 0) stop 2
  st = tracer_get_size(nz, nn)
  allocate(coord(2 * nn), t(nz, nn), s(nz, nn))
  st = tracer_get_mesh_array(cstr("coord"), coord, int(2 * nn, c_long_long))
  do n = 1, nn
    x = coord(2 * n - 1); y = coord(2 * n)
    do k = 1, nz
      zn = real(k - 1, c_double) / real(nz - 1, c_double)
      t(k, n) = 15.0d0 + 5.0d0 * sin(2.0d0 * pi * x / lx) * cos(2.0d0 * pi * y / ly) - 10.0d0 * zn
      s(k, n) = 34.0d0 + (0.5d0 + 0.5d0 * sin(2.0d0 * pi * x / lx)) * (1.0d0 - zn)
    end do
  end do
  st = tracer_allocate_tracers(2_c_int, int(nl, c_int), int(nn, c_int), 1_c_int)
  st = tracer_set_values(1_c_int, t, int(nl, c_int), int(nn, c_int))
  st = tracer_set_values(2_c_int, s, int(nl, c_int), int(nn, c_int))
  do k = 1, 2
    st = tracer_set_schemes(int(k, c_int), cstr("MUSCL"), cstr("QR4C"), cstr("FCT"), 1.0d0, 1.0d0)
    if (st /= 0) stop 3
  end do
  st = tracer_make_velocity(0.1d0, 0.05d0, 0.2d0, 0.5d0, 0.3d0, dt)
  do k = 1, nsteps
    st = tracer_advect_steps(1_c_int, dt)
    if (st /= 0) stop 4
    st = tracer_get_stats(tmin, tmax, tsum)
    write(*, '(A,I4,A,3ES24.16)') 'step ', k, ': T min max sum ', tmin, tmax, tsum
  end do
  st = tracer_get_diagnostics(2_c_int, d)
  write(*, '(A,3ES24.16)') 'S active min max mass ', d(4), d(5), d(6)
  st = tracer_cleanup()
end program analytic_driver
