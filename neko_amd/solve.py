"""Host-side mirror of Neko's solve-path interfaces on top of the C ABI.

Names follow Neko: a `Problem` bundles what Neko keeps in space_t, mesh_t,
dofmap_t, coef_t, gs_t and the zero-Dirichlet bc list; `AxHelm.compute`
mirrors ax_t%compute (src/math/ax.f90:70-84); `Problem.gs_op` mirrors
gs_t%op(u, n, GS_OP_ADD) (src/gs/gather_scatter.f90:1613); `CG.solve`
mirrors ksp_t%solve returning a ksp_monitor_t (src/krylov/krylov.f90:56-66).

Device memory is torch tensors on the current HIP device; every compute
call goes to libneko_amd.so.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import Monitor, OpInfo, check, lib

GS_OP_ADD, GS_OP_MUL, GS_OP_MIN, GS_OP_MAX = 1, 2, 3, 4  # gs_ops.f90:35-36

FIELD_G11, FIELD_G22, FIELD_G33, FIELD_G12, FIELD_G13, FIELD_G23 = range(6)
FIELD_B, FIELD_MULT, FIELD_DINV, FIELD_X, FIELD_Y, FIELD_Z = range(6, 12)
FIELD_H2B, FIELD_JAC = 12, 13


def _ptr(t: torch.Tensor, n: int | None = None, dtype=torch.float64) -> int:
    """Device address of a vector handed to the C ABI, which cannot check
    what it receives: a contiguous device tensor of the expected dtype and
    length (a wrong dtype or size would read and write outside it)."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a device (cuda) tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if n is not None and t.numel() != n:
        raise ValueError(f"expected {n} entries, got {t.numel()}")
    return t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


@dataclass
class KspMonitor:
    """ksp_monitor_t (krylov.f90:56-66) plus the refinement sweep count."""
    iter: int
    res_start: float
    res_final: float
    converged: bool
    outer: int = 0


class Space:
    """GLL space (space_t): nodes, weights, D and D^T (column-major)."""

    def __init__(self, lx: int):
        self.lx = lx
        self.zg = np.zeros(lx)
        self.wx = np.zeros(lx)
        self.dx = np.zeros(lx * lx)
        self.dxt = np.zeros(lx * lx)
        check(lib().nm_gll(lx, self.zg.ctypes.data, self.wx.ctypes.data,
                           self.dx.ctypes.data, self.dxt.ctypes.data), "nm_gll")
        w = self.wx
        self.w3 = np.einsum("k,j,i->kji", w, w, w).reshape(-1).copy()


def box_mesh(nelx, nely, nelz, amp=0.0, seed=0):
    v = np.zeros(nelx * nely * nelz * 24)
    check(lib().nm_box_mesh(nelx, nely, nelz, float(amp), int(seed),
                            v.ctypes.data), "nm_box_mesh")
    return v


def box_gs_map(nelx, nely, nelz, lx, periodic=(0, 0, 0)):
    """Neko's local gs maps of a box (1-based numpy arrays)."""
    L = lib()
    nl, fo, nb = C.c_int64(), C.c_int64(), C.c_int64()
    p = [int(bool(q)) for q in periodic]
    check(L.nm_box_gs_sizes(nelx, nely, nelz, lx, *p, C.byref(nl), C.byref(fo)),
          "nm_box_gs_sizes")
    dg = np.zeros(nl.value, dtype=np.int32)
    gd = np.zeros(nl.value, dtype=np.int32)
    bl = np.zeros(max(fo.value, 1), dtype=np.int32)
    bo = np.zeros(max(fo.value, 1), dtype=np.int32)
    check(L.nm_box_gs_map(nelx, nely, nelz, lx, *p, dg.ctypes.data,
                          gd.ctypes.data, bl.ctypes.data, bo.ctypes.data,
                          C.byref(nl), C.byref(fo), C.byref(nb)), "nm_box_gs_map")
    return dict(dg=dg, gd=gd, blk_len=bl[: nb.value].copy(),
                blk_off=bo[: nb.value].copy(), nlocal=nl.value,
                facet_offset=fo.value, nblks=nb.value)


def box_mask(nelx, nely, nelz, lx, periodic=(0, 0, 0)):
    L = lib()
    p = [int(bool(q)) for q in periodic]
    m = L.nm_box_mask_size(nelx, nely, nelz, lx, *p)
    msk = np.zeros(max(m, 1), dtype=np.int32)
    mm = C.c_int64()
    check(L.nm_box_mask(nelx, nely, nelz, lx, *p, msk.ctypes.data, C.byref(mm)),
          "nm_box_mask")
    return msk[: mm.value].copy()


class Cyclic:
    """Cyclic (rotationally periodic about z) facets of a problem
    (nm_cyclic_create): rotate(vx, vy, vz, idir) is Neko's rotate_cyc."""

    def __init__(self, problem, facets):
        fc = np.ascontiguousarray(np.asarray(facets, dtype=np.int32).reshape(-1, 2))
        h = C.c_void_p()
        check(lib().nm_cyclic_create(C.byref(h), problem.h, fc.ctypes.data, len(fc),
                                     _stream()), "nm_cyclic_create")
        self.h = h
        self.n = problem.n

    def info(self):
        """(cyc_msk 1-based, R11, R12) in entry order."""
        nc = C.c_int64()
        check(lib().nm_cyclic_info(self.h, C.byref(nc), None, None, None), "nm_cyclic_info")
        m = np.zeros(max(nc.value, 1), dtype=np.int32)
        a = np.zeros(max(nc.value, 1))
        b = np.zeros(max(nc.value, 1))
        check(lib().nm_cyclic_info(self.h, C.byref(nc), m.ctypes.data, a.ctypes.data,
                                   b.ctypes.data), "nm_cyclic_info")
        return m[: nc.value], a[: nc.value], b[: nc.value]

    def rotate(self, vx, vy, vz, idir):
        check(lib().nm_rotate_cyc(self.h, _ptr(vx, self.n), _ptr(vy, self.n),
                                  _ptr(vz, self.n), int(idir), _stream()), "nm_rotate_cyc")

    def free(self):
        if self.h:
            lib().nm_cyclic_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fdm_set_variant(variant: int) -> None:
    """Kernel family of the Schwarz FDM solves (nm_fdm_set_variant): 0 the
    matrix cores (default), 1 the vector ALUs."""
    check(lib().nm_fdm_set_variant(int(variant)), "nm_fdm_set_variant")


def read_nmsh(path):
    """Neko .nmsh file -> (vxyz (nel, 8, 3), pid (nel, 8), el_idx (nel,)),
    vertices in Neko's symmetric order, periodic zones applied
    (nm_nmsh_info / nm_nmsh_read)."""
    b = os.fsencode(str(path))
    nel, nz, nper, ncur = (C.c_int64() for _ in range(4))
    check(lib().nm_nmsh_info(b, C.byref(nel), C.byref(nz), C.byref(nper),
                             C.byref(ncur)), "nm_nmsh_info")
    vxyz = np.empty((nel.value, 8, 3), dtype=np.float64)
    pid = np.empty((nel.value, 8), dtype=np.int64)
    el = np.empty(nel.value, dtype=np.int32)
    check(lib().nm_nmsh_read(b, nel.value, vxyz.ctypes.data, pid.ctypes.data,
                             el.ctypes.data), "nm_nmsh_read")
    return vxyz, pid, el


def read_nmsh_curves(path):
    """Curve records of a .nmsh file (nm_nmsh_curves): (el (nc,) 0-based
    element, data (nc, 60) curve_data(5,12), type (nc, 12) edge types)."""
    b = os.fsencode(str(path))
    nel, nz, nper, ncur = (C.c_int64() for _ in range(4))
    check(lib().nm_nmsh_info(b, C.byref(nel), C.byref(nz), C.byref(nper),
                             C.byref(ncur)), "nm_nmsh_info")
    nc = ncur.value
    el = np.zeros(max(nc, 1), dtype=np.int32)
    data = np.zeros((max(nc, 1), 60), dtype=np.float64)
    ty = np.zeros((max(nc, 1), 12), dtype=np.int32)
    check(lib().nm_nmsh_curves(b, nc, el.ctypes.data, data.ctypes.data, ty.ctypes.data),
          "nm_nmsh_curves")
    return el[:nc], data[:nc], ty[:nc]


def curve_xyz(lx, vxyz_e, data, type_):
    """GLL coordinates (3, lx^3) of one element from its 8 vertices and one
    curve record (nm_curve_xyz)."""
    v = np.ascontiguousarray(vxyz_e, dtype=np.float64).reshape(24)
    d = np.ascontiguousarray(data, dtype=np.float64).reshape(60)
    t = np.ascontiguousarray(type_, dtype=np.int32).reshape(12)
    out = np.zeros((3, lx ** 3))
    check(lib().nm_curve_xyz(int(lx), v.ctypes.data, d.ctypes.data, t.ctypes.data,
                             out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data),
          "nm_curve_xyz")
    return out


def read_nmsh_zones(path):
    """Zone records of a .nmsh file as an (nzones, 4) int32 array: element
    (0-based), facet 1..6, type (5 periodic, 7 labelled boundary), label
    (nm_nmsh_zones)."""
    b = os.fsencode(str(path))
    nel, nz, nper, ncur = (C.c_int64() for _ in range(4))
    check(lib().nm_nmsh_info(b, C.byref(nel), C.byref(nz), C.byref(nper),
                             C.byref(ncur)), "nm_nmsh_info")
    out = np.zeros((max(nz.value, 1), 4), dtype=np.int32)
    check(lib().nm_nmsh_zones(b, nz.value, out.ctypes.data), "nm_nmsh_zones")
    return out[: nz.value]


def zone_facets(zones, label=None):
    """(element, facet) pairs of the labelled boundary zones (type 7), in
    file order; label None: all of them (the whole non-periodic boundary)."""
    z = np.asarray(zones)
    sel = z[:, 2] == 7
    if label is not None:
        sel &= z[:, 3] == int(label)
    return [(int(e), int(f)) for e, f in z[sel][:, :2]]


class Problem:
    """One discretised problem resident on the device (geometry, gs maps,
    mask, mass, multiplicity, Jacobi diagonal)."""

    def __init__(self, handle: int, keep=None):
        self.h = C.c_void_p(handle)
        self._keep = keep
        info = OpInfo()
        check(lib().nm_op_info(self.h, C.byref(info)), "nm_op_info")
        self.info = info
        self.n = info.n
        self.lx = info.lx
        self.nel = info.nel
        self.volume = info.volume

    # ---- construction ---------------------------------------------------
    @classmethod
    def box(cls, nelx, nely, nelz, lx, periodic=(0, 0, 0), amp=0.0, seed=0,
            h1=1.0, h2=0.0):
        h = C.c_void_p()
        p = [int(bool(q)) for q in periodic]
        check(lib().nm_op_create_box(C.byref(h), nelx, nely, nelz, lx, *p,
                                     float(amp), int(seed), float(h1),
                                     float(h2), _stream()), "nm_op_create_box")
        pr = cls(h.value)
        pr.box_dims = (nelx, nely, nelz)
        pr.periodic = tuple(p)
        return pr

    @classmethod
    def mesh(cls, vxyz, pid, lx, h1=1.0, h2=0.0, curves=None):
        """General hex mesh (vxyz: (nel, 8, 3) vertices in Neko's symmetric
        order, pid: (nel, 8) point ids), Dirichlet on every facet without a
        neighbour (nm_op_create_mesh); curves = (el, data, type) as
        read_nmsh_curves returns them (nm_op_create_mesh_curved)."""
        v = np.ascontiguousarray(vxyz, dtype=np.float64).reshape(-1, 24)
        p = np.ascontiguousarray(pid, dtype=np.int64).reshape(-1, 8)
        if len(v) != len(p):
            raise ValueError("vxyz and pid disagree on the element count")
        h = C.c_void_p()
        if curves is not None and len(curves[0]):
            el = np.ascontiguousarray(curves[0], dtype=np.int32)
            cd = np.ascontiguousarray(curves[1], dtype=np.float64).reshape(-1, 60)
            ty = np.ascontiguousarray(curves[2], dtype=np.int32).reshape(-1, 12)
            if not (len(el) == len(cd) == len(ty)):
                raise ValueError("curve arrays disagree on the record count")
            check(lib().nm_op_create_mesh_curved(
                C.byref(h), len(v), int(lx), v.ctypes.data, p.ctypes.data, len(el),
                el.ctypes.data, cd.ctypes.data, ty.ctypes.data, float(h1), float(h2),
                _stream()), "nm_op_create_mesh_curved")
        else:
            check(lib().nm_op_create_mesh(C.byref(h), len(v), int(lx), v.ctypes.data,
                                          p.ctypes.data, float(h1), float(h2),
                                          _stream()), "nm_op_create_mesh")
        return cls(h.value)

    @classmethod
    def from_nmsh(cls, path, lx, h1=1.0, h2=0.0):
        """Problem on a Neko .nmsh mesh, curved elements included
        (read_nmsh + read_nmsh_curves + mesh)."""
        vxyz, pid, _ = read_nmsh(path)
        return cls.mesh(vxyz, pid, lx, h1=h1, h2=h2, curves=read_nmsh_curves(path))

    def set_coeffs(self, h1: float, h2: float):
        """New constant Helmholtz coefficients (nm_op_set_coeffs)."""
        check(lib().nm_op_set_coeffs(self.h, float(h1), float(h2), _stream()),
              "nm_op_set_coeffs")

    def opgrad(self, u: torch.Tensor):
        """Weak gradient (ux, uy, uz) of u (Neko's opgrad), local."""
        out = [torch.empty_like(u) for _ in range(3)]
        check(lib().nm_opgrad(self.h, *[_ptr(o, self.n) for o in out], _ptr(u, self.n), _stream()),
              "nm_opgrad")
        return out

    def vel_res(self, u, v, w, p, f, mu, rho, bd, dt):
        """pnpn velocity residual: sets h1 = mu, h2 = rho bd / dt and returns
        (-A u - opgrad_x p + f_x, ...) (nm_vel_res), local."""
        out = [torch.empty_like(u) for _ in range(3)]
        check(lib().nm_vel_res(self.h, float(mu), float(rho), float(bd), float(dt),
                               _ptr(u, self.n), _ptr(v, self.n), _ptr(w, self.n),
                               _ptr(p, self.n), *[_ptr(q, self.n) for q in f],
                               *[_ptr(o, self.n) for o in out],
                               _stream()), "nm_vel_res")
        return out

    def curl(self, u1, u2, u3):
        """Neko's curl: Binv gs(B curl u) (nm_curl)."""
        out = [torch.empty_like(u1) for _ in range(3)]
        check(lib().nm_curl(self.h, *[_ptr(o, self.n) for o in out], _ptr(u1, self.n),
                            _ptr(u2, self.n), _ptr(u3, self.n),
                            _stream()), "nm_curl")
        return out

    def prs_res(self, p, ue, f, mu, rho):
        """pnpn pressure residual without surface terms (nm_prs_res):
        sets h1 = 1/rho, h2 = 0 and returns -A p + sum_c cdtp(ta_c)."""
        out = torch.empty_like(p)
        check(lib().nm_prs_res(self.h, float(mu), float(rho), _ptr(p, self.n),
                               *[_ptr(q, self.n) for q in ue], *[_ptr(q, self.n) for q in f],
                               _ptr(out, self.n),
                               _stream()), "nm_prs_res")
        return out

    def prs_res_surf(self, p, ue, u, f, mu, rho, bd, dt, prs=None, sym=None):
        """The whole pnpn pressure residual (nm_prs_res_surf): prs_res, then
        the facet-normal terms of the velocity-prescribed surface `prs` (with
        the current velocity u) and of the symmetry surface `sym` (Surface
        objects of this problem, or None)."""
        for sf in (prs, sym):
            if sf is not None and sf.problem is not self:
                raise ValueError("surface of another problem")
        out = torch.empty_like(p)
        uu = [_ptr(q, self.n) for q in u] if u is not None else [None, None, None]
        check(lib().nm_prs_res_surf(self.h, float(mu), float(rho), float(bd), float(dt),
                                    _ptr(p, self.n), *[_ptr(q, self.n) for q in ue], *uu,
                                    *[_ptr(q, self.n) for q in f],
                                    prs.h if prs is not None else None,
                                    sym.h if sym is not None else None,
                                    _ptr(out, self.n), _stream()), "nm_prs_res_surf")
        return out

    def set_pc(self, kind: str):
        """Preconditioner of this problem's solvers: "jacobi" (default) or
        "identity" (plain CG, the ksp default without a preconditioner)."""
        k = {"jacobi": 0, "identity": 1}[kind]
        check(lib().nm_op_set_pc(self.h, k, _stream()), "nm_op_set_pc")

    @classmethod
    def from_arrays(cls, nel, lx, dx, g, B, mult_unused, gs, msk, h1=None,
                    h1c=1.0, h2=None, dfrmd=None):
        """From Neko-style host arrays (numpy): g = [G11,G22,G33,G12,G13,G23],
        gs = dict(dg, gd, blk_len, blk_off, nlocal, facet_offset, nblks)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in [dx, *g, B]]
        h1a = None if h1 is None else np.ascontiguousarray(h1, dtype=np.float64)
        h2a = None if h2 is None else np.ascontiguousarray(h2, dtype=np.float64)
        dfa = None if dfrmd is None else np.ascontiguousarray(dfrmd, dtype=np.int32)
        ints = [np.ascontiguousarray(gs[k], dtype=np.int32)
                for k in ("dg", "gd", "blk_len", "blk_off")]
        m = np.ascontiguousarray(msk, dtype=np.int32)
        h = C.c_void_p()
        vp = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        check(lib().nm_op_create(
            C.byref(h), int(nel), int(lx), *[a.ctypes.data for a in arrs],
            vp(h1a), float(h1c), vp(h2a), int(h2a is not None), vp(dfa),
            int(gs["nlocal"]), int(gs["facet_offset"]), int(gs["nblks"]),
            *[a.ctypes.data for a in ints], int(len(m)), m.ctypes.data,
            _stream()), "nm_op_create")
        return cls(h.value, keep=(arrs, h1a, h2a, dfa, ints, m))

    def free(self):
        if self.h and not getattr(self, "_borrowed", False):
            lib().nm_op_free(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def rhs_normal_device(self, seed: int) -> torch.Tensor:
        """Same seeded RHS as rhs_normal, generated on the device
        (nm_op_rhs_normal; equal to it up to last-ulp differences of the
        device log/cos)."""
        f = torch.empty(self.n, dtype=torch.float64, device="cuda")
        check(lib().nm_op_rhs_normal(self.h, int(seed), _ptr(f, self.n), _stream()),
              "nm_op_rhs_normal")
        return f

    # ---- data -----------------------------------------------------------
    def field(self, which: int) -> torch.Tensor:
        """Copy of a device field as a new torch tensor."""
        ptr = lib().nm_op_field(self.h, which)
        if not ptr:
            raise ValueError(f"field {which} not present")
        out = torch.empty(self.n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _memcpy_d2d(out.data_ptr(), ptr, self.n * 8)
        return out

    def gs_maps(self):
        info = self.info
        dg = np.zeros(info.nlocal, dtype=np.int32)
        gd = np.zeros(info.nlocal, dtype=np.int32)
        bl = np.zeros(max(info.nblks, 1), dtype=np.int32)
        bo = np.zeros(max(info.nblks, 1), dtype=np.int32)
        check(lib().nm_op_gs_maps(self.h, dg.ctypes.data, gd.ctypes.data,
                                  bl.ctypes.data, bo.ctypes.data), "gs_maps")
        return dict(dg=dg, gd=gd, blk_len=bl[: info.nblks],
                    blk_off=bo[: info.nblks], nlocal=info.nlocal,
                    facet_offset=info.facet_offset, nblks=info.nblks)

    def set_epb(self, epb64=0, epb32=0):
        check(lib().nm_op_set_epb(self.h, int(epb64), int(epb32)), "set_epb")

    def set_gs_variant(self, v):
        check(lib().nm_op_set_gs_variant(self.h, int(v)), "set_gs_variant")

    def set_cg_face_sums(self, on):
        """Let the fused CG update add the descriptor-face partner values
        (default) or run the face gs pass instead."""
        check(lib().nm_op_set_cg_face_sums(self.h, int(bool(on))), "set_cg_face_sums")
        check(lib().nm_op_info(self.h, C.byref(self.info)), "nm_op_info")

    def autotune(self, reps=5):
        """Pick the fastest Ax / gs kernel variants on this problem."""
        a, b, g = C.c_int(), C.c_int(), C.c_int()
        check(lib().nm_op_autotune(self.h, int(reps), C.byref(a), C.byref(b),
                                   C.byref(g), _stream()), "autotune")
        return {"ax64": a.value, "ax32": b.value, "gs": g.value}

    # ---- operations (Neko names) ----------------------------------------
    def ax_local(self, w: torch.Tensor, u: torch.Tensor):
        """ax_t%compute: w = A_local u (no assembly)."""
        if u.dtype == torch.float32:
            check(lib().nm_op_ax_local32(self.h, _ptr(w, self.n, torch.float32),
                                          _ptr(u, self.n, torch.float32), _stream()),
                  "ax_local32")
        else:
            check(lib().nm_op_ax_local(self.h, _ptr(w, self.n), _ptr(u, self.n), _stream()),
                  "ax_local")
        return w

    def ax_local_vector(self, au, av, aw, u, v, w):
        """ax_t%compute_vector: (au, av, aw) = A_local (u, v, w)."""
        n = self.n
        check(lib().nm_op_ax_local_vector(self.h, _ptr(au, n), _ptr(av, n), _ptr(aw, n),
                                          _ptr(u, n), _ptr(v, n), _ptr(w, n), _stream()),
              "ax_local_vector")
        return au, av, aw

    def ax_gs(self, w, u):
        """w = gs(A_local u): the Ax+gs unit of work."""
        if u.dtype == torch.float32:
            check(lib().nm_op_ax_gs32(self.h, _ptr(w, self.n, torch.float32),
                                       _ptr(u, self.n, torch.float32), _stream()), "ax_gs32")
        else:
            check(lib().nm_op_ax_gs(self.h, _ptr(w, self.n), _ptr(u, self.n), _stream()), "ax_gs")
        return w

    def ax_full(self, w, u):
        """Ax%compute + gs_h%op(ADD) + blst%apply (cg.f90:207-209)."""
        check(lib().nm_op_ax(self.h, _ptr(w, self.n), _ptr(u, self.n), _stream()), "ax")
        return w

    def gs_op(self, u, op=GS_OP_ADD):
        if op != GS_OP_ADD:
            raise NotImplementedError("handle gs is ADD only; use the drop-in "
                                      "hip_gather_kernel for MUL/MIN/MAX")
        check(lib().nm_op_gs(self.h, _ptr(u, self.n), _stream()), "gs")
        return u

    def bc_apply(self, u):
        check(lib().nm_op_mask(self.h, _ptr(u, self.n), _stream()), "mask")
        return u

    def rhs_normal(self, seed: int) -> torch.Tensor:
        """RHS i.i.d. N(0,1) per local dof (seeded), made consistent by one
        gs sum, then masked."""
        f = np.zeros(self.n)
        check(lib().nm_rng_normal_fill(int(seed), self.n, f.ctypes.data), "rng")
        t = torch.from_numpy(f).to("cuda")
        self.gs_op(t)
        self.bc_apply(t)
        return t

    def rhs_sine(self) -> torch.Tensor:
        """b = mask(gs(B * 3 pi^2 sin(pi x) sin(pi y) sin(pi z)))."""
        x, y, z = (self.field(c) for c in (FIELD_X, FIELD_Y, FIELD_Z))
        pi = np.pi
        f = self.field(FIELD_B) * (3 * pi * pi) * torch.sin(pi * x) * \
            torch.sin(pi * y) * torch.sin(pi * z)
        f = f.contiguous()
        self.gs_op(f)
        self.bc_apply(f)
        return f


class CG:
    """Jacobi-preconditioned CG (ksp_t 'cg' with pc_t 'jacobi').
    precision="fp64" or "mixed" (FP32 CG with FP64 residual replacement)."""

    def __init__(self, problem: Problem, precision: str = "fp64"):
        self.problem = problem
        h = C.c_void_p()
        check(lib().nm_cg_create(C.byref(h), problem.h,
                                 0 if precision == "fp64" else 1), "nm_cg_create")
        self.h = h
        self.precision = precision

    def solve(self, x: torch.Tensor, f: torch.Tensor, abs_tol: float,
              max_iter: int = 100000, inner_rel: float = 0.0) -> KspMonitor:
        mon = Monitor()
        n = self.problem.n
        check(lib().nm_cg_solve(self.h, _ptr(x, n), _ptr(f, n), float(abs_tol),
                                int(max_iter), float(inner_rel), C.byref(mon),
                                _stream()), "nm_cg_solve")
        return KspMonitor(mon.iter, mon.res_start, mon.res_final,
                          bool(mon.converged), mon.outer)

    def set_mask(self, msk):
        """This solver's own zero-Dirichlet mask (1-based local indices; one
        velocity component's bc list), or None for the operator's."""
        if msk is None:
            check(lib().nm_cg_set_mask(self.h, None, 0, _stream()), "nm_cg_set_mask")
            return
        m = np.ascontiguousarray(msk, dtype=np.int32)
        check(lib().nm_cg_set_mask(self.h, m.ctypes.data, len(m), _stream()),
              "nm_cg_set_mask")

    def free(self):
        if self.h:
            lib().nm_cg_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def solve_coupled(solvers, xs, fs, abs_tol, max_iter=100000, inner_rel=0.0):
    """ksp_t%solve_coupled of cg (cg.f90:282-305): component c of the
    velocity solved by solvers[c] (same operator, its own mask) -> three
    KspMonitor (nm_cg_solve_coupled)."""
    if len(solvers) != 3 or len(xs) != 3 or len(fs) != 3:
        raise ValueError("solve_coupled takes three solvers, solutions and right-hand sides")
    n = solvers[0].problem.n
    hs = (C.c_void_p * 3)(*[sv.h.value for sv in solvers])
    xp = (C.c_void_p * 3)(*[_ptr(x, n) for x in xs])
    fp = (C.c_void_p * 3)(*[_ptr(f, n) for f in fs])
    mons = (Monitor * 3)()
    check(lib().nm_cg_solve_coupled(hs, xp, fp, float(abs_tol), int(max_iter),
                                    float(inner_rel), mons, _stream()),
          "nm_cg_solve_coupled")
    return [KspMonitor(m.iter, m.res_start, m.res_final, bool(m.converged), m.outer)
            for m in mons]


class Surface:
    """Facet-normal surface of a boundary zone (Neko's facet_normal_t): the
    unique points of the given (element 0-based, facet 1..6) pairs and their
    area-weighted normals (nm_surface_create)."""

    def __init__(self, problem: Problem, facets):
        import numpy as np
        f = np.ascontiguousarray(np.asarray(facets, dtype=np.int32).reshape(-1, 2))
        self.problem = problem
        self.h = C.c_void_p()
        check(lib().nm_surface_create(C.byref(self.h), problem.h,
                                      f.ctypes.data_as(C.c_void_p), int(f.shape[0]),
                                      _stream()), "nm_surface_create")

    def normals(self):
        """(points, nx, ny, nz) as numpy arrays (host copies)."""
        import numpy as np
        npt = C.c_int64()
        check(lib().nm_surface_info(self.h, C.byref(npt)), "nm_surface_info")
        m = npt.value
        pts = np.zeros(max(m, 1), dtype=np.int32)
        nx, ny, nz = (np.zeros(max(m, 1)) for _ in range(3))
        check(lib().nm_surface_normals(self.h, pts.ctypes.data_as(C.c_void_p),
                                       nx.ctypes.data_as(C.c_void_p),
                                       ny.ctypes.data_as(C.c_void_p),
                                       nz.ctypes.data_as(C.c_void_p)), "nm_surface_normals")
        return pts[:m], nx[:m], ny[:m], nz[:m]

    def free(self):
        if self.h:
            lib().nm_surface_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Projection:
    """Solution-space projection around repeated solves (Neko's
    projection_t, src/common/projection.f90): pre() projects the RHS in
    place before a solve, post() adds the projected part back to the
    solution and takes it into the basis (up to L vectors, A-orthonormal).
    dt_last_change < 0: constant time step."""

    def __init__(self, problem: Problem, L: int, activ_step: int = 5):
        self.h = C.c_void_p()
        self.problem = problem
        check(lib().nm_proj_create(C.byref(self.h), problem.h, int(L),
                                   int(activ_step)), "nm_proj_create")

    def pre(self, b: torch.Tensor, tstep: int, dt_last_change: int = -1):
        check(lib().nm_proj_pre(self.h, _ptr(b, self.problem.n), int(tstep), int(dt_last_change),
                                _stream()), "nm_proj_pre")

    def post(self, x: torch.Tensor, tstep: int, dt_last_change: int = -1):
        check(lib().nm_proj_post(self.h, _ptr(x, self.problem.n), int(tstep), int(dt_last_change),
                                 _stream()), "nm_proj_post")

    def clear(self):
        check(lib().nm_proj_clear(self.h, _stream()), "nm_proj_clear")

    def info(self):
        """(basis size m, vectors used by the last projection, its original
        residual sqrt(b.mult.b / volume))"""
        m, pm, res = C.c_int(), C.c_int(), C.c_double()
        check(lib().nm_proj_info(self.h, C.byref(m), C.byref(pm), C.byref(res),
                                 _stream()), "nm_proj_info")
        return m.value, pm.value, res.value

    def free(self):
        if self.h:
            lib().nm_proj_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HSMG:
    """hsmg preconditioner (pc_t 'hsmg', src/krylov/pc_hsmg.f90) on a box
    problem: Schwarz/FDM smoothing on lx and lx_mid, Jacobi-CG on lx=2.
    facet_bc: facet_type of boundary facets (1 pressure, 2 velocity)."""

    def __init__(self, problem: Problem, facet_bc: int = 1, niter: int = 10):
        self.problem = problem
        h = C.c_void_p()
        check(lib().nm_hsmg_create(C.byref(h), problem.h, int(facet_bc),
                                   int(niter), _stream()), "nm_hsmg_create")
        self.h = h

    def apply(self, z: torch.Tensor, r: torch.Tensor):
        """z = M^{-1} r (pc_hsmg.f90:481-627); r is left untouched."""
        check(lib().nm_hsmg_apply(self.h, _ptr(z, self.problem.n), _ptr(r, self.problem.n),
                                   _stream()),
              "nm_hsmg_apply")

    def schwarz_fine(self, e: torch.Tensor, r: torch.Tensor):
        """e = schwarz(r) on the fine level; r is masked in place."""
        check(lib().nm_hsmg_schwarz_fine(self.h, _ptr(e, self.problem.n),
                                          _ptr(r, self.problem.n), _stream()),
              "nm_hsmg_schwarz_fine")

    def free(self):
        if self.h:
            lib().nm_hsmg_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PHMG:
    """phmg-style V-cycle (Chebyshev(smooth_its) smoothing accelerated by
    Schwarz on hsmg's levels, coarse Jacobi-CG) on a box problem.
    level_h2: "default" (Neko's coef default h2 = 1 on the coarser levels)
    or "fine" (the fine level's h2 on every level).  smoother: "fp64", or
    "fp32" (smoothing in FP32, transfers and coarse solve FP64)."""

    def __init__(self, problem: Problem, facet_bc: int = 2, smooth_its: int = 4,
                 level_h2: str = "default", smoother: str = "fp64"):
        self.problem = problem
        h = C.c_void_p()
        check(lib().nm_phmg_create(C.byref(h), problem.h, int(facet_bc),
                                   int(smooth_its), 1 if level_h2 == "fine" else 0,
                                   1 if smoother == "fp32" else 0, _stream()),
              "nm_phmg_create")
        self.h = h

    def apply(self, z: torch.Tensor, r: torch.Tensor):
        check(lib().nm_phmg_apply(self.h, _ptr(z, self.problem.n), _ptr(r, self.problem.n),
                                   _stream()),
              "nm_phmg_apply")

    def set_coarse_amg(self, levels=3, iterations=1, cheby_degree=4):
        """phmg's tree-AMG as the coarse solve (levels 0: back to the
        Jacobi-CG)."""
        check(lib().nm_phmg_set_coarse_amg(self.h, int(levels), int(iterations),
                                           int(cheby_degree), _stream()),
              "nm_phmg_set_coarse_amg")

    def coarse_amg_levels(self):
        nn = (C.c_int64 * 16)()
        k = lib().nm_phmg_coarse_amg_levels(self.h, nn, 16)
        return [nn[i] for i in range(k + 1)] if k else []

    def coarse_amg_map(self, lvl):
        n0 = self.coarse_amg_levels()[0]
        out = np.zeros(n0, dtype=np.int32)
        check(lib().nm_phmg_coarse_amg_map(self.h, int(lvl), out.ctypes.data),
              "nm_phmg_coarse_amg_map")
        return out

    def eigs(self):
        out = (C.c_double * 4)()
        check(lib().nm_phmg_eigs(self.h, out), "nm_phmg_eigs")
        return [(out[0], out[1]), (out[2], out[3])]

    def free(self):
        if self.h:
            lib().nm_phmg_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GMRES:
    """Restarted GMRES(m) with right preconditioning (ksp_t 'gmres',
    src/krylov/bcknd/cpu/gmres.f90).  pc: 'jacobi', 'identity', an HSMG or
    a PHMG."""

    def __init__(self, problem: Problem, pc="jacobi", lgmres: int = 30):
        self.problem = problem
        self.pc = pc
        h = C.c_void_p()
        if isinstance(pc, HSMG):
            code, ph = 1, pc.h
        elif isinstance(pc, PHMG):
            code, ph = 3, pc.h
        elif pc == "jacobi":
            code, ph = 0, None
        elif pc == "identity":
            code, ph = 2, None
        else:
            raise ValueError(f"unknown preconditioner {pc!r}")
        check(lib().nm_gmres_create(C.byref(h), problem.h, code, ph,
                                    int(lgmres)), "nm_gmres_create")
        self.h = h

    def solve(self, x: torch.Tensor, f: torch.Tensor, abs_tol: float,
              max_iter: int = 10000) -> KspMonitor:
        mon = Monitor()
        n = self.problem.n
        check(lib().nm_gmres_solve(self.h, _ptr(x, n), _ptr(f, n), float(abs_tol),
                                   int(max_iter), C.byref(mon), _stream()),
              "nm_gmres_solve")
        return KspMonitor(mon.iter, mon.res_start, mon.res_final,
                          bool(mon.converged), mon.outer)

    def free(self):
        if self.h:
            lib().nm_gmres_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many box problems of uneven size and order solved together (FP32 CG
    with FP64 residual replacement per problem)."""

    def __init__(self, problems):
        self.problems = list(problems)
        arr = (C.c_void_p * len(self.problems))(*[p.h.value for p in self.problems])
        h = C.c_void_p()
        check(lib().nm_batch_create(C.byref(h), len(self.problems), arr, _stream()),
              "nm_batch_create")
        self.h = h

    def autotune(self, reps=3):
        lx = (C.c_int * 16)()
        epb = (C.c_int * 16)()
        ng = C.c_int()
        check(lib().nm_batch_autotune(self.h, int(reps), lx, epb, C.byref(ng)),
              "nm_batch_autotune")
        return {int(lx[g]): int(epb[g]) for g in range(ng.value)}

    def solve(self, xs, fs, rel_tol=1e-10, max_iter=20000, delta=0.0):
        k = len(self.problems)
        xa = (C.c_void_p * k)(*[_ptr(x, p.n) for x, p in zip(xs, self.problems)])
        fa = (C.c_void_p * k)(*[_ptr(f, p.n) for f, p in zip(fs, self.problems)])
        mons = (Monitor * k)()
        check(lib().nm_batch_solve(self.h, xa, fa, float(rel_tol), int(max_iter),
                                   float(delta), mons, _stream()), "nm_batch_solve")
        return [KspMonitor(m.iter, m.res_start, m.res_final, bool(m.converged),
                           m.outer) for m in mons]

    def stats(self):
        """(graph steps, dof-iterations covered by the Ax launches) of the
        last solve (nm_batch_stats)."""
        st, dof = C.c_int64(), C.c_double()
        check(lib().nm_batch_stats(self.h, C.byref(st), C.byref(dof)), "nm_batch_stats")
        return st.value, dof.value

    def info(self):
        n, b = C.c_int64(), C.c_int64()
        check(lib().nm_batch_info(self.h, C.byref(n), C.byref(b)), "nm_batch_info")
        return n.value, b.value

    def free(self):
        if self.h:
            lib().nm_batch_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Serving loop for streams of Poisson solves on resident box meshes
    (no counterpart in Neko).  precision: 'fp64' or 'mixed'."""

    def __init__(self, hbm_budget_bytes: float, max_inflight: int = 8,
                 precision: str = "mixed", chunk_iters: int = 8,
                 keep_solutions: bool = False):
        from ._lib import ServeResult  # noqa: F401  (binding check)
        h = C.c_void_p()
        check(lib().nm_server_create(C.byref(h), float(hbm_budget_bytes),
                                     int(max_inflight),
                                     0 if precision == "fp64" else 1,
                                     int(chunk_iters), int(bool(keep_solutions))),
              "nm_server_create")
        self.h = h
        self.meshes = []

    def add_box_mesh(self, nelx, nely, nelz, lx=8, amp=0.1, seed=0, prewarm=1,
                     tune_reps=0):
        mid = C.c_int()
        check(lib().nm_server_add_box_mesh(self.h, nelx, nely, nelz, lx,
                                           float(amp), int(seed), int(tune_reps),
                                           int(prewarm), C.byref(mid), _stream()),
              "nm_server_add_box_mesh")
        self.meshes.append((nelx, nely, nelz, lx, amp, seed))
        return mid.value

    def add_mesh(self, vxyz, pid, lx=8, prewarm=1, tune_reps=0):
        """Register a general straight-edged mesh (e.g. read_nmsh output)."""
        v = np.ascontiguousarray(vxyz, dtype=np.float64).reshape(-1, 24)
        p = np.ascontiguousarray(pid, dtype=np.int64).reshape(-1, 8)
        mid = C.c_int()
        check(lib().nm_server_add_mesh(self.h, len(v), int(lx), v.ctypes.data, p.ctypes.data,
                                       int(tune_reps), int(prewarm), C.byref(mid),
                                       _stream()), "nm_server_add_mesh")
        self.meshes.append(("mesh", len(v), lx))
        return mid.value

    def mesh_problem(self, mesh_id: int) -> "Problem":
        """The resident operator of a mesh (owned by the server)."""
        oph = C.c_void_p()
        check(lib().nm_server_mesh_op(self.h, int(mesh_id), C.byref(oph)),
              "nm_server_mesh_op")
        pr = Problem.__new__(Problem)
        pr.h = oph
        pr._keep = self
        pr._borrowed = True
        info = OpInfo()
        check(lib().nm_op_info(oph, C.byref(info)), "nm_op_info")
        pr.info, pr.n, pr.lx, pr.nel, pr.volume = (info, info.n, info.lx,
                                                   info.nel, info.volume)
        return pr

    def submit(self, mesh_id, rhs_seed, rel_tol=1e-8, max_iter=20000,
               t_arrive=0.0) -> int:
        rid = C.c_int()
        check(lib().nm_server_submit(self.h, int(mesh_id), int(rhs_seed),
                                     float(rel_tol), int(max_iter),
                                     float(t_arrive), C.byref(rid)),
              "nm_server_submit")
        return rid.value

    def run(self) -> float:
        torch.cuda.synchronize()
        wall = C.c_double()
        check(lib().nm_server_run(self.h, C.byref(wall)), "nm_server_run")
        return wall.value

    def serve(self) -> float:
        """Live stream (nm_server_serve): blocks, serving what other threads
        submit (ctypes releases the GIL for the call), until close() has
        been called and the submitted requests are done; returns the wall
        time."""
        torch.cuda.synchronize()
        wall = C.c_double()
        check(lib().nm_server_serve(self.h, C.byref(wall)), "nm_server_serve")
        return wall.value

    def close(self):
        check(lib().nm_server_close(self.h), "nm_server_close")

    def clock(self) -> float:
        return lib().nm_server_clock(self.h)

    def result(self, rid: int):
        from ._lib import ServeResult
        r = ServeResult()
        check(lib().nm_server_result(self.h, int(rid), C.byref(r)),
              "nm_server_result")
        return r

    def solution(self, rid: int) -> torch.Tensor:
        r = self.result(rid)
        x = torch.empty(int(r.n), dtype=torch.float64, device="cuda")
        check(lib().nm_server_solution(self.h, int(rid), _ptr(x), _stream()),
              "nm_server_solution")
        return x

    def memory(self):
        used, peak = C.c_int64(), C.c_int64()
        check(lib().nm_server_memory(self.h, C.byref(used), C.byref(peak)),
              "nm_server_memory")
        return used.value, peak.value

    def free(self):
        if self.h:
            lib().nm_server_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _memcpy_d2d(dst: int, src: int, nbytes: int):
    check(lib().nm_memcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, 3,
                          _stream()), "nm_memcpy")
