! cathy_gpu_mod.f90 -- ISO_C_BINDING interface of libcathy_gpu.so for the
! CATHY Fortran host.  Every C entry point takes its arguments by reference
! (include/cathy_gpu.h), so the interfaces below pass the host's own
! INTEGER*4 / REAL*8 arrays unchanged, with no copies and no VALUE.
module cathy_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: CATHY_GPU_OK = 0
  integer(c_int), parameter :: CATHY_GPU_LSFAIL = 1
  integer(c_int), parameter :: CATHY_GPU_ORDER_NATURAL = 0
  integer(c_int), parameter :: CATHY_GPU_ORDER_COLUMN3 = 1

  interface
    type(c_ptr) function cathy_gpu_last_error() bind(C, name="cathy_gpu_last_error")
      import :: c_ptr
    end function

    integer(c_int) function cathy_gpu_symslv(n, nterm, numdir, niter, itmxcg, tolcg, rmin, noddir, ja, &
                                              topol, pnew, tnoti, coef1, err, res) &
        bind(C, name="cathy_gpu_symslv")
      import :: c_int, c_double
      integer(c_int), intent(in) :: n, nterm, numdir, itmxcg
      integer(c_int), intent(out) :: niter
      real(c_double), intent(in) :: tolcg, rmin
      integer(c_int), intent(in) :: noddir(*), ja(*), topol(*)
      real(c_double), intent(out) :: pnew(*)
      real(c_double), intent(in) :: tnoti(*), coef1(*)
      real(c_double), intent(out) :: err, res
    end function

    integer(c_int) function cathy_gpu_symslv_release() bind(C, name="cathy_gpu_symslv_release")
      import :: c_int
    end function

    integer(c_int) function cathy_gpu_flow3d_create(handle, n, nt, ntri, nnod, nstr, nzone, maxstr, tetra, &
                                                    bi, ci, di, volu, ivol, volnod, arenod, permx, permy, &
                                                    permz, snodi, pnodi, vgn, vgrmc, vgpsat, kznod, lump, &
                                                    ordering) bind(C, name="cathy_gpu_flow3d_create")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(out) :: handle
      integer(c_int), intent(in) :: n, nt, ntri, nnod, nstr, nzone, maxstr, lump, ordering
      integer(c_int), intent(in) :: tetra(5, *), ivol(*)
      real(c_double), intent(in) :: bi(4, *), ci(4, *), di(4, *), volu(*), volnod(*), arenod(*)
      real(c_double), intent(in) :: permx(maxstr, *), permy(maxstr, *), permz(maxstr, *)
      real(c_double), intent(in) :: snodi(*), pnodi(*), vgn(*), vgrmc(*), vgpsat(*), kznod(*)
    end function

    integer(c_int) function cathy_gpu_flow3d_destroy(handle) bind(C, name="cathy_gpu_flow3d_destroy")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(inout) :: handle
    end function

    integer(c_int) function cathy_gpu_flow3d_set_params(handle, iprm, dprm) &
        bind(C, name="cathy_gpu_flow3d_set_params")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: iprm(9)
      real(c_double), intent(in) :: dprm(16)
    end function

    integer(c_int) function cathy_gpu_flow3d_set_forcing(handle, natm, atmtim, atmval) &
        bind(C, name="cathy_gpu_flow3d_set_forcing")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: natm
      real(c_double), intent(in) :: atmtim(*), atmval(*)
    end function

    ! BCONE/BCNXT/BCBAK record series of the Dirichlet (kind 0: CONTP,
    ! PRESC) or Neumann (kind 1: CONTQ, Q) nodes, expanded as RDNDBC/READBC
    integer(c_int) function cathy_gpu_flow3d_set_bc(handle, kind, nrec, times, counts, nodes, values) &
        bind(C, name="cathy_gpu_flow3d_set_bc")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: kind, nrec
      real(c_double), intent(in) :: times(*), values(*)
      integer(c_int), intent(in) :: counts(*), nodes(*)
    end function

    ! ETRAN's vegetation: VEG_TYPE(NNOD) and the soil file's PCANA .. OMGC
    integer(c_int) function cathy_gpu_flow3d_set_vegetation(handle, nveg, veg_type, pcana, pcref, pcwlt, &
        zroot, pz, omgc, z) bind(C, name="cathy_gpu_flow3d_set_vegetation")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: nveg, veg_type(*)
      real(c_double), intent(in) :: pcana(*), pcref(*), pcwlt(*), zroot(*), pz(*), omgc(*), z(*)
    end function

    ! seepage faces: NSF faces of NSFNUM(i) nodes, NSFNOD packed face by face
    integer(c_int) function cathy_gpu_flow3d_set_seepage(handle, nsf, counts, nodes, isfcvg) &
        bind(C, name="cathy_gpu_flow3d_set_seepage")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: nsf, counts(*), nodes(*), isfcvg
    end function

    integer(c_int) function cathy_gpu_flow3d_get_seepage(handle, sfex, sfq) &
        bind(C, name="cathy_gpu_flow3d_get_seepage")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(out) :: sfex(*)
      real(c_double), intent(out) :: sfq(*)
    end function

    ! SURF = 1: the coupled switching of ISIMGR = 2 (SWITCH in CONVER, ADRSTN)
    integer(c_int) function cathy_gpu_flow3d_set_surface(handle, surf) &
        bind(C, name="cathy_gpu_flow3d_set_surface")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: surf
    end function

    ! the routing block of a coupled run: a route handle and DATIN's cell
    ! numbering (CELL(5,NCELL_COARSE), INDEX, INDEX_WITH_LAKES, DEM_MAP,
    ! TRIANG(4,NTRI)); the flow step then runs SURF_FLOWTRA and PONDUPD
    integer(c_int) function cathy_gpu_flow3d_set_routing(handle, route, nrow, ncol, delta_x, delta_y, &
        ncell_coarse, cell, indcel, indcelwl, dem_map, ntri, triang) &
        bind(C, name="cathy_gpu_flow3d_set_routing")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle, route
      integer(c_int), intent(in) :: nrow, ncol, ncell_coarse, ntri
      real(c_double), intent(in) :: delta_x, delta_y
      integer(c_int), intent(in) :: cell(*), indcel(*), indcelwl(*), triang(*)
      real(c_double), intent(in) :: dem_map(*)
    end function

    ! initial ponding heads PONDNOD(NNOD) (IPOND 1/2): the run starts ponded
    integer(c_int) function cathy_gpu_flow3d_set_initial_ponding(handle, pondnod) &
        bind(C, name="cathy_gpu_flow3d_set_initial_ponding")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      real(c_double), intent(in) :: pondnod(*)
    end function

    ! PONDNOD(NNOD), OVFLNOD(NNOD), H_WATER_KKP1_SN(NROW*NCOL); INFO(2)
    integer(c_int) function cathy_gpu_flow3d_routing_state(handle, pondnod, ovflnod, h_water, info) &
        bind(C, name="cathy_gpu_flow3d_routing_state")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      real(c_double), intent(out) :: pondnod(*), ovflnod(*), h_water(*)
      integer(c_int), intent(out) :: info(2)
    end function

    integer(c_int) function cathy_gpu_flow3d_init(handle, ptimep) bind(C, name="cathy_gpu_flow3d_init")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      real(c_double), intent(in) :: ptimep(*)
    end function

    integer(c_int) function cathy_gpu_flow3d_advance(handle, stats) bind(C, name="cathy_gpu_flow3d_advance")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      real(c_double), intent(out) :: stats(32)
    end function

    ! PNEW, SW, CKRW, IFATM: pass c_loc(array) or c_null_ptr
    integer(c_int) function cathy_gpu_flow3d_get_state(handle, pnew, sw, ckrw, ifatm) &
        bind(C, name="cathy_gpu_flow3d_get_state")
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), intent(in) :: handle
      type(c_ptr), value :: pnew, sw, ckrw, ifatm
    end function

    integer(c_int) function cathy_gpu_flow3d_set_pnew(handle, pnew) bind(C, name="cathy_gpu_flow3d_set_pnew")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      real(c_double), intent(in) :: pnew(*)
    end function

    integer(c_int) function cathy_gpu_flow3d_counters(handle, out6) bind(C, name="cathy_gpu_flow3d_counters")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      real(c_double), intent(out) :: out6(6)
    end function

    ! surface routing substep (ROUTE + ALTEZZE), route.f:4-234, altezze.f:4-60
    integer(c_int) function cathy_gpu_route_create(handle, nrow, ncol, ldc, ncell, qoi_sn, p_outflow_1, &
                                                   p_outflow_2, w_1, w_2, local_slope_1, local_slope_2, &
                                                   epl_1, epl_2, kss1_sf_1, kss1_sf_2, ws1_sf_1, ws1_sf_2, &
                                                   b1_sf, y1_sf, nrc) bind(C, name="cathy_gpu_route_create")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(out) :: handle
      integer(c_int), intent(in) :: nrow, ncol, ldc, ncell, qoi_sn(*), p_outflow_1(*), p_outflow_2(*)
      real(c_double), intent(in) :: w_1(*), w_2(*), local_slope_1(*), local_slope_2(*), epl_1(*), epl_2(*)
      real(c_double), intent(in) :: kss1_sf_1(*), kss1_sf_2(*), ws1_sf_1(*), ws1_sf_2(*), b1_sf(*)
      real(c_double), intent(in) :: y1_sf(*), nrc(*)
    end function

    integer(c_int) function cathy_gpu_route_step(handle, nmember, deltat, delta_x, delta_y, &
                                                 surface_water_sn, q_in_kk_sn, q_out_kk_sn_1, q_out_kk_sn_2, &
                                                 volume_kk_sn, q_in_kkp1_sn, q_out_kkp1_sn_1, &
                                                 q_out_kkp1_sn_2, volume_kkp1_sn, h_water_kkp1_sn, cu_ak, &
                                                 h_pool_kk_vec, h_pool_kkp1_vec) &
        bind(C, name="cathy_gpu_route_step")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: nmember
      real(c_double), intent(in) :: deltat, delta_x, delta_y
      real(c_double), intent(in) :: surface_water_sn(*), q_in_kk_sn(*), q_out_kk_sn_1(*), q_out_kk_sn_2(*)
      real(c_double), intent(in) :: volume_kk_sn(*)
      real(c_double), intent(inout) :: q_in_kkp1_sn(*), q_out_kkp1_sn_1(*), q_out_kkp1_sn_2(*)
      real(c_double), intent(inout) :: volume_kkp1_sn(*)
      real(c_double), intent(out) :: h_water_kkp1_sn(*)
      real(c_double), intent(inout) :: cu_ak(2, *)
      real(c_double), intent(in) :: h_pool_kk_vec(*)
      real(c_double), intent(inout) :: h_pool_kkp1_vec(*)
    end function

    integer(c_int) function cathy_gpu_route_set_reservoirs(handle, tipo_r, reservr, numres, maxres, n_ha, &
                                                           hres, ares, h_soglia, h_fondo, elevation) &
        bind(C, name="cathy_gpu_route_set_reservoirs")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: tipo_r(*), reservr(*), numres, maxres, n_ha(*)
      real(c_double), intent(in) :: hres(maxres, *), ares(maxres, *), h_soglia(*), h_fondo(*), elevation(*)
    end function

    integer(c_int) function cathy_gpu_route_set_celtype(handle, celtype) bind(C, name="cathy_gpu_route_set_celtype")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(in) :: handle
      integer(c_int), intent(in) :: celtype(*)
    end function

    ! lockstep ensemble of flow handles (one launch per kernel for all members)
    integer(c_int) function cathy_gpu_ensemble_create(ens, nmember, members) &
        bind(C, name="cathy_gpu_ensemble_create")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(out) :: ens
      integer(c_int), intent(in) :: nmember
      integer(c_int64_t), intent(in) :: members(*)
    end function

    integer(c_int) function cathy_gpu_ensemble_advance(ens, nsteps, rc, taken, stats) &
        bind(C, name="cathy_gpu_ensemble_advance")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), intent(in) :: ens
      integer(c_int), intent(in) :: nsteps
      integer(c_int), intent(out) :: rc(*), taken(*)
      real(c_double), intent(out) :: stats(32, nsteps, *)
    end function

    integer(c_int) function cathy_gpu_ensemble_set_graphs(ens, on) bind(C, name="cathy_gpu_ensemble_set_graphs")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(in) :: ens
      integer(c_int), intent(in) :: on
    end function

    integer(c_int) function cathy_gpu_ensemble_destroy(ens) bind(C, name="cathy_gpu_ensemble_destroy")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(inout) :: ens
    end function

    integer(c_int) function cathy_gpu_route_destroy(handle) bind(C, name="cathy_gpu_route_destroy")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(inout) :: handle
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Text of the library's last error on this thread.
  function cathy_gpu_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, l
    p = cathy_gpu_last_error()
    l = int(c_strlen(p))
    call c_f_pointer(p, chars, [l])
    allocate(character(len=l) :: msg)
    do i = 1, l
      msg(i:i) = chars(i)
    end do
  end function

end module cathy_gpu
