! neko_amd_capi -- iso_c_binding interfaces of libneko_amd.so
! (include/neko_amd.h) for Neko's Fortran: the drop-in entry points with
! the conventions Neko's own HIP interfaces use (device pointers by value,
! integers by reference: src/math/bcknd/device/ax_helm_device.F90:53-64,
! src/gs/bcknd/device/gs_device.F90, src/math/bcknd/device/device_math.F90,
! src/krylov/bcknd/device/pc_jacobi_device.F90) and the handle API that a
! registered ksp_t / pc_t plug-in calls (register_krylov,
! src/krylov/krylov.f90:260; register_precon,
! src/krylov/precon_fctry.f90:123-156).  Needs only iso_c_binding, so it
! compiles and links without the rest of Neko.
module neko_amd_capi
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_int, c_int64_t, c_double, c_ptr, c_char, c_size_t

  type, bind(c) :: nm_monitor_t
     integer(c_int) :: iter, outer
     real(c_double) :: res_start, res_final
     integer(c_int) :: converged
  end type nm_monitor_t

  type, bind(c) :: nm_gmres_monitor_t
     integer(c_int) :: iter, outer
     real(c_double) :: res_start, res_final
     integer(c_int) :: converged
  end type nm_gmres_monitor_t

  interface
     ! ---- drop-in entry points (same names and conventions as Neko's HIP)
     subroutine hip_ax_helm(w, u, dx, dy, dz, dxt, dyt, dzt, h1, g11, g22, &
          g33, g12, g13, g23, nelv, lx) bind(c, name='hip_ax_helm')
       import c_ptr, c_int
       type(c_ptr), value :: w, u, dx, dy, dz, dxt, dyt, dzt, h1, &
            g11, g22, g33, g12, g13, g23
       integer(c_int) :: nelv, lx
     end subroutine hip_ax_helm

     subroutine hip_gather_kernel(v, m, o, dg, u, n, gd, nb, b, bo, op, &
          strm) bind(c, name='hip_gather_kernel')
       import c_ptr, c_int
       integer(c_int) :: m, o, n, nb, op
       type(c_ptr), value :: v, dg, u, gd, b, bo, strm
     end subroutine hip_gather_kernel

     subroutine hip_scatter_kernel(v, m, dg, u, n, gd, nb, b, bo, strm) &
          bind(c, name='hip_scatter_kernel')
       import c_ptr, c_int
       integer(c_int) :: m, n, nb
       type(c_ptr), value :: v, dg, u, gd, b, bo, strm
     end subroutine hip_scatter_kernel

     real(c_double) function hip_glsc3(a, b, c, n, strm) &
          bind(c, name='hip_glsc3')
       import c_ptr, c_int, c_double
       type(c_ptr), value :: a, b, c, strm
       integer(c_int) :: n
     end function hip_glsc3

     subroutine hip_glsc3_many(h, w, v, mult, j, n, strm) &
          bind(c, name='hip_glsc3_many')
       import c_ptr, c_int, c_double
       real(c_double) :: h(*)
       type(c_ptr), value :: w, v, mult, strm
       integer(c_int) :: j, n
     end subroutine hip_glsc3_many

     subroutine hip_add2s1(a, b, c1, n, strm) bind(c, name='hip_add2s1')
       import c_ptr, c_int, c_double
       type(c_ptr), value :: a, b, strm
       real(c_double) :: c1
       integer(c_int) :: n
     end subroutine hip_add2s1

     subroutine hip_add2s2(a, b, c1, n, strm) bind(c, name='hip_add2s2')
       import c_ptr, c_int, c_double
       type(c_ptr), value :: a, b, strm
       real(c_double) :: c1
       integer(c_int) :: n
     end subroutine hip_add2s2

     subroutine hip_col3(a, b, c, n, strm) bind(c, name='hip_col3')
       import c_ptr, c_int
       type(c_ptr), value :: a, b, c, strm
       integer(c_int) :: n
     end subroutine hip_col3

     subroutine hip_jacobi_update(d, dxt, dyt, dzt, g11, g22, g33, g12, &
          g13, g23, nel, lx) bind(c, name='hip_jacobi_update')
       import c_ptr, c_int
       type(c_ptr), value :: d, dxt, dyt, dzt, g11, g22, g33, g12, g13, g23
       integer(c_int) :: nel, lx
     end subroutine hip_jacobi_update

     subroutine hip_rotate_cyc(vx, vy, vz, x, y, z, cyc_msk, r11, r12, &
          ncyc, idir) bind(c, name='hip_rotate_cyc')
       import c_ptr, c_int
       type(c_ptr), value :: vx, vy, vz, x, y, z, cyc_msk, r11, r12
       integer(c_int) :: ncyc, idir
     end subroutine hip_rotate_cyc

     ! ---- handle API
     function nm_last_error() bind(c, name='nm_last_error')
       import c_ptr
       type(c_ptr) :: nm_last_error
     end function nm_last_error

     integer(c_int) function nm_gll(lx, z, w, d, dt) bind(c, name='nm_gll')
       import c_int, c_double
       integer(c_int), value :: lx
       real(c_double) :: z(*), w(*), d(*), dt(*)
     end function nm_gll

     integer(c_int) function nm_box_gs_sizes(nelx, nely, nelz, lx, perx, &
          pery, perz, nlocal, facet_offset) bind(c, name='nm_box_gs_sizes')
       import c_int, c_int64_t
       integer(c_int), value :: nelx, nely, nelz, lx, perx, pery, perz
       integer(c_int64_t) :: nlocal, facet_offset
     end function nm_box_gs_sizes

     integer(c_int) function nm_box_gs_map(nelx, nely, nelz, lx, perx, &
          pery, perz, dg, gd, blk_len, blk_off, nlocal, facet_offset, nblks) &
          bind(c, name='nm_box_gs_map')
       import c_int, c_int64_t
       integer(c_int), value :: nelx, nely, nelz, lx, perx, pery, perz
       integer(c_int) :: dg(*), gd(*), blk_len(*), blk_off(*)
       integer(c_int64_t) :: nlocal, facet_offset, nblks
     end function nm_box_gs_map

     integer(c_int) function nm_op_create(op, nel, lx, dx, g11, g22, g33, &
          g12, g13, g23, B, h1, h1c, h2, ifh2, dfrmd, nlocal, facet_offset, &
          nblks, dg, gd, blk_len, blk_off, nmsk, msk, s) &
          bind(c, name='nm_op_create')
       import c_int, c_ptr, c_double
       type(c_ptr) :: op
       integer(c_int), value :: nel, lx, ifh2, nlocal, facet_offset, nblks, nmsk
       type(c_ptr), value :: dx, g11, g22, g33, g12, g13, g23, B, h1, h2, &
            dfrmd, dg, gd, blk_len, blk_off, msk, s
       real(c_double), value :: h1c
     end function nm_op_create

     subroutine nm_op_free(op) bind(c, name='nm_op_free')
       import c_ptr
       type(c_ptr), value :: op
     end subroutine nm_op_free

     integer(c_int) function nm_op_set_coeffs(op, h1, h2, s) &
          bind(c, name='nm_op_set_coeffs')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: op, s
       real(c_double), value :: h1, h2
     end function nm_op_set_coeffs

     integer(c_int) function nm_cg_create(cg, op, precision) &
          bind(c, name='nm_cg_create')
       import c_int, c_ptr
       type(c_ptr) :: cg
       type(c_ptr), value :: op
       integer(c_int), value :: precision
     end function nm_cg_create

     subroutine nm_cg_free(cg) bind(c, name='nm_cg_free')
       import c_ptr
       type(c_ptr), value :: cg
     end subroutine nm_cg_free

     integer(c_int) function nm_cg_solve(cg, x, f, abs_tol, max_iter, &
          inner_rel, mon, s) bind(c, name='nm_cg_solve')
       import c_int, c_ptr, c_double, nm_monitor_t
       type(c_ptr), value :: cg, x, f, s
       real(c_double), value :: abs_tol, inner_rel
       integer(c_int), value :: max_iter
       type(nm_monitor_t) :: mon
     end function nm_cg_solve

     integer(c_int) function nm_hsmg_create(h, op, facet_bc, niter, s) &
          bind(c, name='nm_hsmg_create')
       import c_int, c_ptr
       type(c_ptr) :: h
       type(c_ptr), value :: op, s
       integer(c_int), value :: facet_bc, niter
     end function nm_hsmg_create

     integer(c_int) function nm_hsmg_apply(h, z, r, s) &
          bind(c, name='nm_hsmg_apply')
       import c_int, c_ptr
       type(c_ptr), value :: h, z, r, s
     end function nm_hsmg_apply

     integer(c_int) function nm_gmres_create(g, op, pc, pc_handle, lgmres) &
          bind(c, name='nm_gmres_create')
       import c_int, c_ptr
       type(c_ptr) :: g
       type(c_ptr), value :: op, pc_handle
       integer(c_int), value :: pc, lgmres
     end function nm_gmres_create

     integer(c_int) function nm_gmres_solve(g, x, f, abs_tol, max_iter, mon, &
          s) bind(c, name='nm_gmres_solve')
       import c_int, c_ptr, c_double, nm_gmres_monitor_t
       type(c_ptr), value :: g, x, f, s
       real(c_double), value :: abs_tol
       integer(c_int), value :: max_iter
       type(nm_gmres_monitor_t) :: mon
     end function nm_gmres_solve

     ! ---- velocity solve: a solver's own zero-Dirichlet list (a velocity
     ! component's bc list blstx / blsty / blstz, 1-based), solve_coupled
     ! (src/krylov/bcknd/cpu/cg.f90:282-305) and compute_vector
     ! (src/math/ax.f90:98-115)
     integer(c_int) function nm_cg_set_mask(cg, msk, nmsk, s) &
          bind(c, name='nm_cg_set_mask')
       import c_int, c_int64_t, c_ptr
       type(c_ptr), value :: cg, s
       integer(c_int), intent(in) :: msk(*)
       integer(c_int64_t), value :: nmsk
     end function nm_cg_set_mask

     integer(c_int) function nm_cg_solve_coupled(cg, x, f, abs_tol, max_iter, &
          inner_rel, mon, s) bind(c, name='nm_cg_solve_coupled')
       import c_int, c_ptr, c_double, nm_monitor_t
       type(c_ptr), intent(in) :: cg(3), x(3), f(3)
       real(c_double), value :: abs_tol, inner_rel
       integer(c_int), value :: max_iter
       type(nm_monitor_t) :: mon(3)
       type(c_ptr), value :: s
     end function nm_cg_solve_coupled

     integer(c_int) function nm_op_ax_local_vector(op, au, av, aw, u, v, w, s) &
          bind(c, name='nm_op_ax_local_vector')
       import c_int, c_ptr
       type(c_ptr), value :: op, au, av, aw, u, v, w, s
     end function nm_op_ax_local_vector

     ! ---- phmg (Chebyshev/Schwarz p-multigrid) and its tree-AMG coarse
     ! solve (src/multigrid/phmg.f90:138-187; levels, V-cycles and the
     ! Chebyshev degree as phmg's json keys; levels 0 = hsmg's coarse CG)
     integer(c_int) function nm_phmg_set_coarse_amg(h, levels, iterations, &
          cheby_degree, s) bind(c, name='nm_phmg_set_coarse_amg')
       import c_int, c_ptr
       type(c_ptr), value :: h, s
       integer(c_int), value :: levels, iterations, cheby_degree
     end function nm_phmg_set_coarse_amg
  end interface
end module neko_amd_capi
