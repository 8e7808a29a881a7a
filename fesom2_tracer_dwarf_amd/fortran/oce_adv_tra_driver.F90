!> FESOM2's tracer-advection step, do_oce_adv_tra, on the GPU, for FESOM host
!! code: a drop-in for the module of lib/oce_adv_tra_driver.F90.
!!
!! The module name, the routine and its argument list are the reference's
!! (do_oce_adv_tra(dt, vel, w, wi, we, tr_num, dynamics, tracers, partit,
!! mesh), lib/oce_adv_tra_driver.F90:23-45), and the module is compiled
!! against the host's own MOD_MESH, MOD_TRACER, MOD_PARTIT and MOD_DYN, so a
!! host that says "use oce_adv_tra_driver_module" links this one instead.
!!
!! The first call hands the mesh to the library (through tracer_mesh_begin /
!! tracer_set_mesh_array / tracer_mesh_end; after fta_mesh_changed, the next
!! call hands it over again). Every call then passes the velocity (vel, w,
!! we), the tracer's values and valuesAB and its schemes, runs the whole step
!! on the GPU and ADDS the horizontal and vertical tendencies into
!! tracers%work%del_ttf_advhoriz and del_ttf_advvert, as the reference routine
!! does; the tracer itself is not changed. The MUSCL prerequisites are formed
!! on the device from valuesAB (the nodal and up/down-wind gradients of
!! fill_up_dn_grad), so the host's edge_up_dn_grad is not read; its
!! edge_up_dn_tri and nboundary_lay are used when allocated.
!!
!! One partition (eDim_nod2D = 0, as on one GPU), no cavities (ulevels = 1),
!! no implicit vertical split (use_wsplit = .false.); anything else stops
!! with an error, as the reference stops on an unknown scheme.
module oce_adv_tra_driver_module
  use iso_c_binding
  use O_PARAM, only: WP, MP
  use MOD_MESH
  use MOD_TRACER
  use MOD_PARTIT
  use MOD_DYN
  use fesom_tracer_amd
  implicit none
  private
  public :: do_oce_adv_tra, fta_mesh_changed

  logical, save :: mesh_on_device = .false.

contains

  !> the mesh (or its layer thicknesses) changed: hand it over again on the next call
  subroutine fta_mesh_changed()
    mesh_on_device = .false.
  end subroutine fta_mesh_changed

  subroutine check(st, what)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: what
    if (st /= 0) then
      write(*, '(A)') '### error: '//what//': '//last_error_message()
      error stop 1
    end if
  end subroutine check

  !> an index array as the library takes it: 0-based, -1 for "no element" (FESOM's 0)
  subroutine put_index(name, a)
    character(len=*), intent(in) :: name
    integer, intent(in) :: a(:, :)
    integer(c_int), allocatable :: b(:, :)
    allocate(b(size(a, 1), size(a, 2)))
    b = int(a, c_int) - 1_c_int
    where (b < 0) b = -1_c_int
    call check(tracer_set_mesh_array(cstr(name), b, int(size(b), c_long_long)), 'mesh array '//name)
  end subroutine put_index

  subroutine put_int(name, a)
    character(len=*), intent(in) :: name
    integer, intent(in) :: a(:)
    integer(c_int), allocatable :: b(:)
    allocate(b(size(a)))
    b = int(a, c_int)
    call check(tracer_set_mesh_array(cstr(name), b, int(size(b), c_long_long)), 'mesh array '//name)
  end subroutine put_int

  subroutine put_real1(name, a)
    character(len=*), intent(in) :: name
    real(kind=MP), intent(in) :: a(:)
    real(c_double), allocatable :: b(:)
    allocate(b(size(a)))
    b = real(a, c_double)
    call check(tracer_set_mesh_array(cstr(name), b, int(size(b), c_long_long)), 'mesh array '//name)
  end subroutine put_real1

  subroutine put_real2(name, a)
    character(len=*), intent(in) :: name
    real(kind=MP), intent(in) :: a(:, :)
    real(c_double), allocatable :: b(:, :)
    allocate(b(size(a, 1), size(a, 2)))
    b = real(a, c_double)
    call check(tracer_set_mesh_array(cstr(name), b, int(size(b), c_long_long)), 'mesh array '//name)
  end subroutine put_real2

  subroutine hand_over_mesh(tracers, partit, mesh)
    type(t_tracer), intent(in) :: tracers
    type(t_partit), intent(in) :: partit
    type(t_mesh), intent(in) :: mesh
    integer :: n2, e2, t2
    if (partit%eDim_nod2D /= 0) then
      write(*, '(A)') '### error: do_oce_adv_tra on the GPU runs one partition (eDim_nod2D must be 0)'
      error stop 1
    end if
    n2 = mesh%nod2D
    e2 = mesh%edge2D
    t2 = mesh%elem2D
    call check(tracer_mesh_begin(int(mesh%nl, c_int), int(n2, c_int), int(t2, c_int), int(e2, c_int), &
                                 int(size(mesh%nod_in_elem2D, 1), c_int)), 'tracer_mesh_begin')
    call put_index('edges', mesh%edges(:, 1:e2))
    call put_index('edge_tri', mesh%edge_tri(:, 1:e2))
    call put_index('elem2D_nodes', mesh%elem2D_nodes(1:3, 1:t2))
    call put_index('nod_in_elem2D', mesh%nod_in_elem2D(:, 1:n2))
    call put_int('nod_in_elem2D_num', mesh%nod_in_elem2D_num(1:n2))
    call put_int('nlevels', mesh%nlevels(1:t2))
    call put_int('nlevels_nod2D', mesh%nlevels_nod2D(1:n2))
    if (allocated(mesh%ulevels)) call put_int('ulevels', mesh%ulevels(1:t2))
    if (allocated(mesh%ulevels_nod2D)) call put_int('ulevels_nod2D', mesh%ulevels_nod2D(1:n2))
    if (allocated(mesh%nlevels_nod2D_min)) call put_int('nlevels_nod2D_min', mesh%nlevels_nod2D_min(1:n2))
    if (allocated(mesh%ulevels_nod2D_max)) call put_int('ulevels_nod2D_max', mesh%ulevels_nod2D_max(1:n2))
    if (allocated(tracers%work%nboundary_lay)) call put_int('nboundary_lay', tracers%work%nboundary_lay(1:n2))
    if (allocated(tracers%work%edge_up_dn_tri)) call put_index('edge_up_dn_tri', tracers%work%edge_up_dn_tri(:, 1:e2))
    call put_real1('zbar', mesh%zbar)
    call put_real1('elem_cos', mesh%elem_cos(1:t2))
    call put_real1('elem_area', mesh%elem_area(1:t2))
    call put_real2('gradient_sca', mesh%gradient_sca(:, 1:t2))
    call put_real2('edge_dxdy', mesh%edge_dxdy(:, 1:e2))
    call put_real2('edge_cross_dxdy', mesh%edge_cross_dxdy(:, 1:e2))
    call put_real2('area', mesh%area(:, 1:n2))
    if (allocated(mesh%areasvol)) call put_real2('areasvol', mesh%areasvol(:, 1:n2))
    call put_real2('hnode', mesh%hnode(:, 1:n2))
    if (allocated(mesh%hnode_new)) call put_real2('hnode_new', mesh%hnode_new(:, 1:n2))
    call put_real2('helem', mesh%helem(:, 1:t2))
    if (allocated(mesh%Z_3d_n)) call put_real2('Z_3d_n', mesh%Z_3d_n(:, 1:n2))
    if (allocated(mesh%zbar_3d_n)) call put_real2('zbar_3d_n', mesh%zbar_3d_n(:, 1:n2))
    call check(tracer_mesh_end(), 'tracer_mesh_end')
    call check(tracer_allocate_tracers(int(tracers%num_tracers, c_int), int(mesh%nl, c_int), int(n2, c_int), &
                                       1_c_int), 'tracer_allocate_tracers')
    ! valuesAB comes from the host with every call
    call check(tracer_set_option(cstr('refresh_valuesAB'), 0_c_int), 'tracer_set_option')
    mesh_on_device = .true.
  end subroutine hand_over_mesh

  !> lib/oce_adv_tra_driver.F90:23-294, on the GPU
  subroutine do_oce_adv_tra(dt, vel, w, wi, we, tr_num, dynamics, tracers, partit, mesh)
    real(kind=WP),  intent(in),    target :: dt
    integer,        intent(in)            :: tr_num
    type(t_mesh)  , intent(in)   , target :: mesh
    type(t_partit), intent(inout), target :: partit
    type(t_tracer), intent(inout), target :: tracers
    type(t_dyn)   , intent(inout), target :: dynamics
    real(kind=WP),  intent(in)            :: vel(2, mesh%nl-1, partit%myDim_elem2D+partit%eDim_elem2D)
    real(kind=WP),  intent(in), target    :: W(mesh%nl,    partit%myDim_nod2D+partit%eDim_nod2D)
    real(kind=WP),  intent(in), target    :: WI(mesh%nl,   partit%myDim_nod2D+partit%eDim_nod2D)
    real(kind=WP),  intent(in), target    :: WE(mesh%nl,   partit%myDim_nod2D+partit%eDim_nod2D)
    integer(c_int) :: nl, n2, t2, id
    real(c_double), allocatable :: dh(:, :), dv(:, :)

    if (dynamics%use_wsplit) then
      write(*, '(A)') '### error: do_oce_adv_tra on the GPU: the implicit vertical split (use_wsplit) is not supported'
      error stop 1
    end if
    if (.not. mesh_on_device) call hand_over_mesh(tracers, partit, mesh)
    nl = int(mesh%nl, c_int)
    n2 = int(mesh%nod2D, c_int)
    t2 = int(mesh%elem2D, c_int)
    id = int(tr_num, c_int)
    call check(tracer_set_velocity_elem(real(vel(:, :, 1:t2), c_double), real(W(:, 1:n2), c_double), &
                                        real(WE(:, 1:n2), c_double), nl, n2, t2), 'tracer_set_velocity_elem')
    associate (td => tracers%data(tr_num))
      call check(tracer_set_values(id, real(td%values(:, 1:n2), c_double), nl, n2), 'tracer_set_values')
      call check(tracer_set_values_ab(id, real(td%valuesAB(:, 1:n2), c_double), nl, n2), 'tracer_set_values_ab')
      call check(tracer_set_schemes(id, cstr(td%tra_adv_hor), cstr(td%tra_adv_ver), cstr(td%tra_adv_lim), &
                                    real(td%tra_adv_ph, c_double), real(td%tra_adv_pv, c_double)), &
                 'tracer_set_schemes')
    end associate
    allocate(dh(nl - 1, n2), dv(nl - 1, n2))
    dh = real(tracers%work%del_ttf_advhoriz(:, 1:n2), c_double)
    dv = real(tracers%work%del_ttf_advvert(:, 1:n2), c_double)
    call check(tracer_do_oce_adv_tra(id, real(dt, c_double), dh, dv, nl, n2), 'tracer_do_oce_adv_tra')
    tracers%work%del_ttf_advhoriz(:, 1:n2) = real(dh, MP)
    tracers%work%del_ttf_advvert(:, 1:n2) = real(dv, MP)
  end subroutine do_oce_adv_tra

end module oce_adv_tra_driver_module
