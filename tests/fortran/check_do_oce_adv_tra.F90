!> A FESOM-style host of the drop-in do_oce_adv_tra, for the GPU test.
!!
!!   check_do_oce_adv_tra <dir>
!!
!! Reads <dir>/mesh.bin and <dir>/state.bin (stream files the test writes:
!! dimensions, then FESOM's mesh arrays with 1-based indices and 0 for "no
!! element", then the velocity and each tracer's values, valuesAB, schemes
!! and tra_adv_ph/pv), fills t_mesh / t_partit / t_tracer / t_dyn, and for
!! every tracer zeroes del_ttf_advhoriz / del_ttf_advvert, calls
!! do_oce_adv_tra twice (the second call adds onto the first) and writes both
!! results to <dir>/tend.bin.
program check_do_oce_adv_tra
  use O_PARAM
  use MOD_MESH
  use MOD_TRACER
  use MOD_PARTIT
  use MOD_DYN
  use oce_adv_tra_driver_module, only: do_oce_adv_tra
  implicit none
  type(t_mesh) :: mesh
  type(t_partit) :: partit
  type(t_tracer) :: tracers
  type(t_dyn) :: dyn
  character(len=4096) :: dir
  integer :: u, nl, N, T, E, A, ntr, k, pass
  real(kind=WP) :: dt

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/mesh.bin', access='stream', form='unformatted', status='old')
  read(u) nl, N, T, E, A
  mesh%nl = nl; mesh%nod2D = N; mesh%elem2D = T; mesh%edge2D = E
  allocate(mesh%edges(2, E), mesh%edge_tri(2, E), mesh%elem2D_nodes(3, T), mesh%nlevels(T), mesh%ulevels(T))
  allocate(mesh%nlevels_nod2D(N), mesh%ulevels_nod2D(N), mesh%nlevels_nod2D_min(N), mesh%ulevels_nod2D_max(N))
  allocate(mesh%nod_in_elem2D_num(N), mesh%nod_in_elem2D(A, N))
  allocate(tracers%work%edge_up_dn_tri(2, E), tracers%work%nboundary_lay(N))
  read(u) mesh%edges, mesh%edge_tri, mesh%elem2D_nodes, mesh%nlevels, mesh%ulevels, mesh%nlevels_nod2D, &
          mesh%ulevels_nod2D, mesh%nlevels_nod2D_min, mesh%ulevels_nod2D_max, mesh%nod_in_elem2D_num, &
          mesh%nod_in_elem2D, tracers%work%edge_up_dn_tri, tracers%work%nboundary_lay
  allocate(mesh%zbar(nl), mesh%elem_cos(T), mesh%elem_area(T), mesh%gradient_sca(6, T), mesh%edge_dxdy(2, E))
  allocate(mesh%edge_cross_dxdy(4, E), mesh%area(nl, N), mesh%areasvol(nl, N), mesh%hnode(nl - 1, N))
  allocate(mesh%hnode_new(nl - 1, N), mesh%helem(nl - 1, T), mesh%Z_3d_n(nl - 1, N), mesh%zbar_3d_n(nl, N))
  read(u) mesh%zbar, mesh%elem_cos, mesh%elem_area, mesh%gradient_sca, mesh%edge_dxdy, mesh%edge_cross_dxdy, &
          mesh%area, mesh%areasvol, mesh%hnode, mesh%hnode_new, mesh%helem, mesh%Z_3d_n, mesh%zbar_3d_n
  close(u)
  partit%myDim_nod2D = N; partit%eDim_nod2D = 0
  partit%myDim_elem2D = T; partit%eDim_elem2D = 0
  partit%myDim_edge2D = E; partit%eDim_edge2D = 0

  open(newunit=u, file=trim(dir)//'/state.bin', access='stream', form='unformatted', status='old')
  read(u) ntr, dt
  allocate(dyn%uv(2, nl - 1, T), dyn%w(nl, N), dyn%w_e(nl, N), dyn%w_i(nl, N))
  read(u) dyn%uv, dyn%w, dyn%w_e
  dyn%w_i = 0.0_WP
  tracers%num_tracers = ntr
  allocate(tracers%data(ntr))
  do k = 1, ntr
    allocate(tracers%data(k)%values(nl - 1, N), tracers%data(k)%valuesAB(nl - 1, N))
    read(u) tracers%data(k)%values, tracers%data(k)%valuesAB, tracers%data(k)%tra_adv_hor, &
            tracers%data(k)%tra_adv_ver, tracers%data(k)%tra_adv_lim, tracers%data(k)%tra_adv_ph, &
            tracers%data(k)%tra_adv_pv
  end do
  close(u)

  allocate(tracers%work%del_ttf_advhoriz(nl - 1, N), tracers%work%del_ttf_advvert(nl - 1, N))
  open(newunit=u, file=trim(dir)//'/tend.bin', access='stream', form='unformatted', status='replace')
  do k = 1, ntr
    tracers%work%del_ttf_advhoriz = 0.0_MP
    tracers%work%del_ttf_advvert = 0.0_MP
    do pass = 1, 2
      call do_oce_adv_tra(dt, dyn%uv, dyn%w, dyn%w_i, dyn%w_e, k, dyn, tracers, partit, mesh)
      write(u) tracers%work%del_ttf_advhoriz, tracers%work%del_ttf_advvert
    end do
  end do
  close(u)
  write(*, '(A, I0, A)') 'do_oce_adv_tra: ', ntr, ' tracers done'
end program check_do_oce_adv_tra
