"""ctypes binding of libneko_amd.so (the C ABI of include/neko_amd.h).

The library is loaded from the package directory only.  There is no CPU
fallback: if the shared library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libneko_amd.so")

_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pi64 = C.POINTER(C.c_int64)


class OpInfo(C.Structure):
    _fields_ = [("nel", _i64), ("n", _i64), ("nlocal", _i64), ("nblks", _i64),
                ("npairs", _i64), ("nmsk", _i64), ("facet_offset", _i64),
                ("lx", _i), ("volume", _d), ("h1c", _d), ("ifh2", _i),
                ("nfaces", _i64), ("nrest", _i64), ("cg_face_sums", _i),
                ("mult_counts", _i)]


class ServeResult(C.Structure):
    _fields_ = [("state", _i), ("iter", _i), ("outer", _i), ("converged", _i),
                ("res_start", _d), ("res_final", _d), ("t_arrive", _d),
                ("t_admit", _d), ("t_done", _d), ("n", C.c_int64)]


class Monitor(C.Structure):
    _fields_ = [("iter", _i), ("outer", _i), ("res_start", _d),
                ("res_final", _d), ("converged", _i)]


# name: (restype, argtypes)
SIGNATURES = {
    # drop-in entry points
    "hip_ax_helm": (None, [_vp] * 15 + [_pi, _pi]),
    "hip_ax_helm_sp": (None, [_vp] * 10 + [_pi, _pi, _vp]),
    "hip_gather_kernel": (None, [_vp, _pi, _pi, _vp, _vp, _pi, _vp, _pi, _vp,
                                 _vp, _pi, _vp]),
    "hip_scatter_kernel": (None, [_vp, _pi, _vp, _vp, _pi, _vp, _pi, _vp, _vp,
                                  _vp]),
    "hip_glsc3": (_d, [_vp, _vp, _vp, _pi, _vp]),
    "hip_glsc3_many": (None, [C.POINTER(_d), _vp, _vp, _vp, _pi, _pi, _vp]),
    "hip_add2s1": (None, [_vp, _vp, C.POINTER(_d), _pi, _vp]),
    "hip_add2s2": (None, [_vp, _vp, C.POINTER(_d), _pi, _vp]),
    "hip_cmult2": (None, [_vp, _vp, C.POINTER(_d), _pi, _vp]),
    "hip_copy": (None, [_vp, _vp, _pi, _vp]),
    "hip_rzero": (None, [_vp, _pi, _vp]),
    "hip_cmult": (None, [_vp, C.POINTER(_d), _pi, _vp]),
    "hip_cfill": (None, [_vp, C.POINTER(_d), _pi, _vp]),
    "hip_add2": (None, [_vp, _vp, _pi, _vp]),
    "hip_col2": (None, [_vp, _vp, _pi, _vp]),
    "hip_col3": (None, [_vp, _vp, _vp, _pi, _vp]),
    "hip_sub2": (None, [_vp, _vp, _pi, _vp]),
    "hip_invcol1": (None, [_vp, _pi, _vp]),
    "hip_addcol3": (None, [_vp, _vp, _vp, _pi, _vp]),
    "hip_addcol4": (None, [_vp, _vp, _vp, _vp, _pi, _vp]),
    "hip_jacobi_update": (None, [_vp] * 10 + [_pi, _pi]),
    "nm_jacobi_update": (None, [_vp] * 9 + [_pi, _pi, _vp]),
    "nm_glsc3_many_noalloc": (_i, [_vp, _vp, _vp, _vp, _i, _i, _vp]),
    # handle API
    "nm_last_error": (C.c_char_p, []),
    "nm_gll": (_i, [_i, _vp, _vp, _vp, _vp]),
    "nm_rng_uniform": (_d, [C.c_uint64, C.c_uint64]),
    "nm_rng_normal_fill": (_i, [C.c_uint64, _i64, _vp]),
    "nm_box_mesh": (_i, [_i, _i, _i, _d, C.c_uint64, _vp]),
    "nm_dfrmd": (_i, [_i64, _vp, _vp]),
    "nm_xyz": (_i, [_i64, _i, _vp, _vp, _vp, _vp, _vp]),
    "nm_box_gs_sizes": (_i, [_i] * 7 + [_pi64, _pi64]),
    "nm_box_gs_map": (_i, [_i] * 7 + [_vp, _vp, _vp, _vp, _pi64, _pi64, _pi64]),
    "nm_box_mask_size": (_i64, [_i] * 7),
    "nm_box_mask": (_i, [_i] * 7 + [_vp, _pi64]),
    "nm_op_create_box": (_i, [C.POINTER(_vp)] + [_i] * 7 + [_d, C.c_uint64,
                                                          _d, _d, _vp]),
    "nm_nmsh_info": (_i, [C.c_char_p, _pi64, _pi64, _pi64, _pi64]),
    "nm_nmsh_read": (_i, [C.c_char_p, _i64, _vp, _vp, _vp]),
    "nm_nmsh_zones": (_i, [C.c_char_p, _i64, _vp]),
    "nm_nmsh_curves": (_i, [C.c_char_p, _i64, _vp, _vp, _vp]),
    "nm_curve_xyz": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nm_mesh_facet_neigh": (_i, [_i64, _vp, _vp]),
    "nm_mesh_gs_sizes": (_i, [_i64, _i, _vp, _pi64, _pi64]),
    "nm_mesh_gs_map": (_i, [_i64, _i, _vp, _vp, _vp, _vp, _vp, _vp, _pi64,
                            _pi64, _pi64]),
    "nm_mesh_mask_size": (_i64, [_i64, _i, _vp]),
    "nm_mesh_mask": (_i, [_i64, _i, _vp, _vp, _pi64]),
    "nm_op_create_mesh": (_i, [C.POINTER(_vp), _i64, _i, _vp, _vp, _d, _d,
                               _vp]),
    "nm_op_create_mesh_curved": (_i, [C.POINTER(_vp), _i64, _i, _vp, _vp, _i64, _vp,
                                      _vp, _vp, _d, _d, _vp]),
    "nm_op_set_pc": (_i, [_vp, _i, _vp]),
    "nm_op_set_coeffs": (_i, [_vp, _d, _d, _vp]),
    "nm_opgrad": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "nm_vel_res": (_i, [_vp, _d, _d, _d, _d] + [_vp] * 10 + [_vp]),
    "nm_curl": (_i, [_vp] * 8),
    "nm_prs_res": (_i, [_vp, _d, _d] + [_vp] * 9),
    "nm_prs_res_surf": (_i, [_vp, _d, _d, _d, _d] + [_vp] * 14),
    "nm_surface_create": (_i, [C.POINTER(_vp), _vp, _vp, _i64, _vp]),
    "nm_surface_free": (None, [_vp]),
    "nm_surface_info": (_i, [_vp, _pi64]),
    "nm_surface_normals": (_i, [_vp] * 5),
    "nm_op_create": (_i, [C.POINTER(_vp), _i, _i] + [_vp] * 9 +
                     [_d, _vp, _i, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i,
                      _vp, _vp]),
    "nm_cyclic_create": (_i, [C.POINTER(_vp), _vp, _vp, _i64, _vp]),
    "nm_cyclic_free": (None, [_vp]),
    "nm_cyclic_info": (_i, [_vp, _pi64, _vp, _vp, _vp]),
    "nm_rotate_cyc": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "hip_rotate_cyc": (None, [_vp] * 9 + [C.POINTER(_i), C.POINTER(_i)]),
    "nm_op_free": (None, [_vp]),
    "nm_op_info": (_i, [_vp, C.POINTER(OpInfo)]),
    "nm_op_field": (_vp, [_vp, _i]),
    "nm_op_gs_maps": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "nm_op_set_epb": (_i, [_vp, _i, _i]),
    "nm_op_set_gs_variant": (_i, [_vp, _i]),
    "nm_op_set_cg_face_sums": (_i, [_vp, _i]),
    "nm_op_autotune": (_i, [_vp, _i, _pi, _pi, _pi, _vp]),
    "nm_op_ax_local": (_i, [_vp, _vp, _vp, _vp]),
    "nm_op_ax_gs": (_i, [_vp, _vp, _vp, _vp]),
    "nm_op_ax": (_i, [_vp, _vp, _vp, _vp]),
    "nm_op_ax_local32": (_i, [_vp, _vp, _vp, _vp]),
    "nm_op_ax_gs32": (_i, [_vp, _vp, _vp, _vp]),
    "nm_op_gs": (_i, [_vp, _vp, _vp]),
    "nm_op_mask": (_i, [_vp, _vp, _vp]),
    "nm_op_set_all_deformed": (_i, [_vp, _vp]),
    "nm_op_rhs_normal": (_i, [_vp, C.c_uint64, _vp, _vp]),
    "nm_fdm_set_variant": (_i, [_i]),
    "nm_hsmg_create": (_i, [C.POINTER(_vp), _vp, _i, _i, _vp]),
    "nm_hsmg_free": (None, [_vp]),
    "nm_hsmg_apply": (_i, [_vp, _vp, _vp, _vp]),
    "nm_hsmg_schwarz_fine": (_i, [_vp, _vp, _vp, _vp]),
    "nm_gmres_create": (_i, [C.POINTER(_vp), _vp, _i, _vp, _i]),
    "nm_phmg_create": (_i, [C.POINTER(_vp), _vp, _i, _i, _i, _i, _vp]),
    "nm_phmg_free": (None, [_vp]),
    "nm_phmg_apply": (_i, [_vp, _vp, _vp, _vp]),
    "nm_phmg_eigs": (_i, [_vp, C.POINTER(_d)]),
    "nm_gmres_free": (None, [_vp]),
    "nm_gmres_solve": (_i, [_vp, _vp, _vp, _d, _i, C.POINTER(Monitor), _vp]),
    "nm_batch_create": (_i, [C.POINTER(_vp), _i, C.POINTER(_vp), _vp]),
    "nm_batch_free": (None, [_vp]),
    "nm_batch_autotune": (_i, [_vp, _i, _pi, _pi, _pi]),
    "nm_batch_solve": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), _d, _i, _d,
                            C.POINTER(Monitor), _vp]),
    "nm_batch_stats": (_i, [_vp, C.POINTER(C.c_int64), C.POINTER(_d)]),
    "nm_batch_info": (_i, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "nm_server_create": (_i, [C.POINTER(_vp), _d, _i, _i, _i, _i]),
    "nm_server_free": (None, [_vp]),
    "nm_server_add_box_mesh": (_i, [_vp, _i, _i, _i, _i, _d, C.c_uint64, _i, _i,
                                    C.POINTER(_i), _vp]),
    "nm_server_add_mesh": (_i, [_vp, _i64, _i, _vp, _vp, _i, _i, C.POINTER(_i), _vp]),
    "nm_server_mesh_op": (_i, [_vp, _i, C.POINTER(_vp)]),
    "nm_server_submit": (_i, [_vp, _i, C.c_uint64, _d, _i, _d, C.POINTER(_i)]),
    "nm_server_run": (_i, [_vp, C.POINTER(_d)]),
    "nm_server_serve": (_i, [_vp, C.POINTER(_d)]),
    "nm_server_close": (_i, [_vp]),
    "nm_server_clock": (_d, [_vp]),
    "nm_server_result": (_i, [_vp, _i, C.POINTER(ServeResult)]),
    "nm_server_solution": (_i, [_vp, _i, _vp, _vp]),
    "nm_server_memory": (_i, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "nm_proj_create": (_i, [C.POINTER(_vp), _vp, _i, _i]),
    "nm_proj_free": (None, [_vp]),
    "nm_proj_clear": (_i, [_vp, _vp]),
    "nm_proj_pre": (_i, [_vp, _vp, _i, _i, _vp]),
    "nm_proj_post": (_i, [_vp, _vp, _i, _i, _vp]),
    "nm_proj_info": (_i, [_vp, _pi, _pi, C.POINTER(_d), _vp]),
    "nm_phmg_set_coarse_amg": (_i, [_vp, _i, _i, _i, _vp]),
    "nm_phmg_coarse_amg_levels": (_i, [_vp, C.POINTER(C.c_int64), _i]),
    "nm_phmg_coarse_amg_map": (_i, [_vp, _i, _vp]),
    "nm_cg_create": (_i, [C.POINTER(_vp), _vp, _i]),
    "nm_cg_free": (None, [_vp]),
    "nm_cg_solve": (_i, [_vp, _vp, _vp, _d, _i, _d, C.POINTER(Monitor), _vp]),
    "nm_cg_set_mask": (_i, [_vp, _vp, _i64, _vp]),
    "nm_cg_solve_coupled": (_i, [_vp, _vp, _vp, _d, _i, _d, C.POINTER(Monitor), _vp]),
    "nm_op_ax_local_vector": (_i, [_vp] * 7 + [_vp]),
    "nm_memcpy": (_i, [_vp, _vp, C.c_size_t, _i, _vp]),
}

_lib = None


class NekoAmdError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libneko_amd.so (built in-tree by neko_amd.build); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NekoAmdError(
            f"{LIB_PATH} is missing: build it with `python -m neko_amd.build` "
            "(there is no CPU fallback)")
    # torch ships its own HIP runtime (soname libamdhip64.so.7); importing it
    # first makes this library bind to that same runtime instance.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().nm_last_error().decode(errors="replace")
        raise NekoAmdError(f"{what} failed (status {status}): {msg}")
