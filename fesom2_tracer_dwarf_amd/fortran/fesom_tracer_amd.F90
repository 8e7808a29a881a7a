!> ISO_C_BINDING interface of libfesom_tracer_amd.so for Fortran host code.
!!
!! The first group binds the entry points of the reference's C interface
!! (lib/tracer_c_interface.F90) under their own names and argument lists, so a
!! host that called them keeps calling the same routines; the second group
!! binds this library's extensions (mesh generation, schemes, velocity,
!! multi-step advection, diagnostics). Character arguments are C strings:
!! pass trim(s)//c_null_char. Tracer arrays are (nl-1, nod2D) column-major.
module fesom_tracer_amd
  use iso_c_binding
  implicit none
  private :: c_char, c_int, c_double, c_long_long, c_ptr

  interface
    ! ---- entry points of lib/tracer_c_interface.F90
    integer(c_int) function tracer_init_mpi(comm) bind(C, name="tracer_init_mpi")
      import :: c_int
      integer(c_int), value :: comm
    end function
    integer(c_int) function tracer_finalize_mpi() bind(C, name="tracer_finalize_mpi")
      import :: c_int
    end function
    integer(c_int) function tracer_init(restart_dir) bind(C, name="tracer_init")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: restart_dir(*)
    end function
    integer(c_int) function tracer_advect_step(dt) bind(C, name="tracer_advect_step")
      import :: c_int, c_double
      real(c_double), value :: dt
    end function
    integer(c_int) function tracer_get_stats(tmin, tmax, tsum) bind(C, name="tracer_get_stats")
      import :: c_int, c_double
      real(c_double), intent(out) :: tmin, tmax, tsum
    end function
    integer(c_int) function tracer_get_size(nz, nn) bind(C, name="tracer_get_size")
      import :: c_int
      integer(c_int), intent(out) :: nz, nn
    end function
    integer(c_int) function tracer_get_values(values, nz, nn) bind(C, name="tracer_get_values")
      import :: c_int, c_double
      integer(c_int), value :: nz, nn
      real(c_double), intent(out) :: values(nz, nn)
    end function
    integer(c_int) function tracer_finalize(output_dir) bind(C, name="tracer_finalize")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: output_dir(*)
    end function
    integer(c_int) function tracer_cleanup() bind(C, name="tracer_cleanup")
      import :: c_int
    end function
    integer(c_int) function tracer_allocate_tracers(num_tracers, nl, nod2d, ab_order) &
        bind(C, name="tracer_allocate_tracers")
      import :: c_int
      integer(c_int), value :: num_tracers, nl, nod2d, ab_order
    end function
    integer(c_int) function tracer_set_values(tracer_id, values, nl, nod2d) bind(C, name="tracer_set_values")
      import :: c_int, c_double
      integer(c_int), value :: tracer_id, nl, nod2d
      real(c_double), intent(in) :: values(nl - 1, nod2d)
    end function
    integer(c_int) function tracer_set_velocity_zero(nl, nod2d, edge2d) bind(C, name="tracer_set_velocity_zero")
      import :: c_int
      integer(c_int), value :: nl, nod2d, edge2d
    end function
    integer(c_int) function tracer_init_complete() bind(C, name="tracer_init_complete")
      import :: c_int
    end function
    integer(c_int) function tracer_load_mesh_from_files(mesh_path) bind(C, name="tracer_load_mesh_from_files")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: mesh_path(*)
    end function
    integer(c_int) function tracer_init_for_advection() bind(C, name="tracer_init_for_advection")
      import :: c_int
    end function

    ! ---- extensions of this library
    integer(c_int) function tracer_generate_mesh(nx, ny, periodic, lx, ly, nl, dz, stretch, lev_lo, lev_hi, seed, &
                                                 sfc) bind(C, name="tracer_generate_mesh")
      import :: c_int, c_double, c_long_long
      integer(c_int), value :: nx, ny, periodic, nl, lev_lo, lev_hi, sfc
      real(c_double), value :: lx, ly, dz, stretch
      integer(c_long_long), value :: seed
    end function
    integer(c_int) function tracer_set_schemes(tracer_id, hor, ver, lim, ph, pv) bind(C, name="tracer_set_schemes")
      import :: c_int, c_char, c_double
      integer(c_int), value :: tracer_id
      character(kind=c_char), intent(in) :: hor(*), ver(*), lim(*)
      real(c_double), value :: ph, pv
    end function
    integer(c_int) function tracer_make_velocity(u_bg, v_bg, cell_amp, shear, cfl, dt) &
        bind(C, name="tracer_make_velocity")
      import :: c_int, c_double
      real(c_double), value :: u_bg, v_bg, cell_amp, shear, cfl
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function tracer_get_mesh_array(name, out, count) bind(C, name="tracer_get_mesh_array")
      import :: c_int, c_char, c_long_long
      character(kind=c_char), intent(in) :: name(*)
      type(*), dimension(*) :: out
      integer(c_long_long), value :: count
    end function
    integer(c_int) function tracer_get_values_id(tracer_id, values, nz, nn) bind(C, name="tracer_get_values_id")
      import :: c_int, c_double
      integer(c_int), value :: tracer_id, nz, nn
      real(c_double), intent(out) :: values(nz, nn)
    end function
    integer(c_int) function tracer_set_option(name, value) bind(C, name="tracer_set_option")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), value :: value
    end function
    integer(c_int) function tracer_advect_steps(nsteps, dt) bind(C, name="tracer_advect_steps")
      import :: c_int, c_double
      integer(c_int), value :: nsteps
      real(c_double), value :: dt
    end function
    integer(c_int) function tracer_get_diagnostics(tracer_id, out) bind(C, name="tracer_get_diagnostics")
      import :: c_int, c_double
      integer(c_int), value :: tracer_id
      real(c_double), intent(out) :: out(7)
    end function
    integer(c_int) function tracer_mesh_begin(nl, nod2d, elem2d, edge2d, maxadj) bind(C, name="tracer_mesh_begin")
      import :: c_int
      integer(c_int), value :: nl, nod2d, elem2d, edge2d, maxadj
    end function
    integer(c_int) function tracer_set_mesh_array(name, data, count) bind(C, name="tracer_set_mesh_array")
      import :: c_int, c_char, c_long_long
      character(kind=c_char), intent(in) :: name(*)
      type(*), dimension(*), intent(in) :: data
      integer(c_long_long), value :: count
    end function
    integer(c_int) function tracer_mesh_end() bind(C, name="tracer_mesh_end")
      import :: c_int
    end function
    integer(c_int) function tracer_set_velocity_elem(uv, w, w_e, nl, nod2d, elem2d) &
        bind(C, name="tracer_set_velocity_elem")
      import :: c_int, c_double
      integer(c_int), value :: nl, nod2d, elem2d
      real(c_double), intent(in) :: uv(2, nl - 1, elem2d), w(nl, nod2d), w_e(nl, nod2d)
    end function
    integer(c_int) function tracer_set_values_ab(tracer_id, values, nl, nod2d) bind(C, name="tracer_set_values_ab")
      import :: c_int, c_double
      integer(c_int), value :: tracer_id, nl, nod2d
      real(c_double), intent(in) :: values(nl - 1, nod2d)
    end function
    integer(c_int) function tracer_do_oce_adv_tra(tracer_id, dt, dttf_h, dttf_v, nl, nod2d) &
        bind(C, name="tracer_do_oce_adv_tra")
      import :: c_int, c_double
      integer(c_int), value :: tracer_id, nl, nod2d
      real(c_double), value :: dt
      real(c_double), intent(inout) :: dttf_h(nl - 1, nod2d), dttf_v(nl - 1, nod2d)
    end function
    type(c_ptr) function tracer_last_error() bind(C, name="tracer_last_error")
      import :: c_ptr
    end function
  end interface

contains

  !> C string from a Fortran string
  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s)//c_null_char
  end function cstr

  !> the library's last error message as a Fortran string
  function last_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    cp = tracer_last_error()
    msg = ""
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function last_error_message

end module fesom_tracer_amd
