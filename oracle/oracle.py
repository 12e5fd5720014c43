"""ctypes front end of the CPU FP64 oracle (oracle/neko_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package (neko_amd) never does.

`OracleCase` builds a whole problem the way Neko's setup does (box mesh ->
facet connectivity -> dofmap -> gs maps -> coef -> mask -> Jacobi) and
exposes Neko's solve-path operations on numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libneko_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


_REF_SO = os.path.join(_HERE, "_ref", "libref_gs_sort.so")
_REF_SRC = "/root/reference/src/gs/gather_scatter.f90"


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def build_ref() -> str | None:
    """Compile the reference's own gs_qsort_dofmap (Fortran, flang) into
    oracle/_ref/ when the reference tree is present (oracle/build_ref.sh);
    None when it is not."""
    if not os.path.exists(_REF_SRC):
        return None
    subprocess.run([os.path.join(_HERE, "build_ref.sh")], check=True,
                   stdout=subprocess.DEVNULL)
    return _REF_SO


def ref_lib():
    """The reference-built gs sort (oracle/_ref), or None if not built."""
    if not os.path.exists(_REF_SO):
        return None
    R = C.CDLL(_REF_SO)
    R.ref_gs_qsort_dofmap.argtypes = [_ip, _ip, C.c_int]
    return R


_lib = None


class OrProblem(C.Structure):
    _fields_ = [
        ("nel", C.c_int), ("lx", C.c_int),
        ("d", C.c_void_p), ("dt", C.c_void_p),
        ("h1", C.c_void_p), ("h2", C.c_void_p), ("B", C.c_void_p),
        ("ifh2", C.c_int),
        ("g11", C.c_void_p), ("g22", C.c_void_p), ("g33", C.c_void_p),
        ("g12", C.c_void_p), ("g13", C.c_void_p), ("g23", C.c_void_p),
        ("mult", C.c_void_p),
        ("gs_m", C.c_int), ("gs_o", C.c_int), ("gs_nb", C.c_int),
        ("dg", C.c_void_p), ("gd", C.c_void_p),
        ("blk_len", C.c_void_p), ("blk_off", C.c_void_p),
        ("gs_buf", C.c_void_p),
        ("msk", C.c_void_p), ("nmsk", C.c_int),
        ("volume", C.c_double),
    ]


class OrMonitor(C.Structure):
    _fields_ = [("iter", C.c_int), ("res_start", C.c_double),
                ("res_final", C.c_double), ("converged", C.c_int)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        build()
    L = C.CDLL(_SO)
    L.or_rng_uniform.restype = C.c_double
    L.or_rng_uniform.argtypes = [C.c_uint64, C.c_uint64]
    L.or_rng_normal.restype = C.c_double
    L.or_rng_normal.argtypes = [C.c_uint64, C.c_uint64]
    L.or_zwgll.argtypes = [_dp, _dp, C.c_int]
    L.or_dgll.argtypes = [_dp, _dp, _dp, C.c_int]
    L.or_box_mesh.argtypes = [C.c_int] * 6 + [C.c_double, C.c_uint64, _dp, _lp]
    L.or_facet_neigh.argtypes = [C.c_int, _lp, _ip]
    L.or_dfrmd_el.argtypes = [C.c_int, _dp, _ip]
    L.or_dof_ids.argtypes = [C.c_int, C.c_int, _lp, _lp]
    L.or_xyz.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.or_curve_element.restype = C.c_int
    L.or_curve_element.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _dp, _dp, _dp]
    L.or_xyz_curved.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _ip,
                                _dp, _dp, _dp, C.POINTER(C.c_int)]
    L.or_gs_register.restype = C.c_int
    L.or_gs_register.argtypes = [C.c_int, C.c_int, _lp, _ip, _ip, _ip,
                                 C.POINTER(C.c_int)]
    L.or_gs_qsort_dofmap.argtypes = [_ip, _ip, C.c_int]
    L.or_gs_map.argtypes = [C.c_int, C.c_int, _lp, _ip, _ip, _ip,
                            C.POINTER(C.c_int), C.POINTER(C.c_int), _ip, _ip,
                            C.POINTER(C.c_int)]
    L.or_gs_op_add.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _ip, _ip,
                               C.c_int, _ip, _ip, _dp]
    L.or_coef.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp] + [_dp] * 8
    L.or_ax_helm.argtypes = ([_dp, _dp, _dp, _dp, _dp] + [_dp] * 6 +
                             [C.c_int, C.c_int, _dp, _dp, C.c_int])
    L.or_mask.restype = C.c_int
    L.or_mask.argtypes = [C.c_int, C.c_int, _ip, C.c_void_p]
    L.or_mask_apply.argtypes = [_dp, _ip, C.c_int]
    L.or_glsc3.restype = C.c_double
    L.or_glsc3.argtypes = [_dp, _dp, _dp, C.c_int64]
    L.or_jacobi_diag.argtypes = [_dp, _dp] + [_dp] * 6 + [_ip, C.c_int, C.c_int]
    L.or_jacobi_diag_dev.argtypes = [_dp, _dp] + [_dp] * 6 + [C.c_int, C.c_int]
    L.or_cg_history.argtypes = [C.c_void_p, C.c_int]
    L.or_cg_solve.restype = OrMonitor
    L.or_cg_solve.argtypes = [C.POINTER(OrProblem), C.c_void_p, _dp, _dp,
                              C.c_int, C.c_double]
    L.or_gmres_solve.restype = OrMonitor
    L.or_gmres_solve.argtypes = [C.POINTER(OrProblem), C.c_void_p, _dp, _dp,
                                 C.c_int, C.c_double, C.c_int]
    L.or_ax_full.argtypes = [C.POINTER(OrProblem), _dp, _dp]
    L.or_cofactors.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.or_opgrad.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.or_dudxyz.argtypes = [C.c_int, C.c_int] + [_dp] * 7
    L.or_cdtp.argtypes = [C.c_int, C.c_int] + [_dp] * 7
    _lib = L
    return L


def zwgll(lx):
    z = np.zeros(lx)
    w = np.zeros(lx)
    lib().or_zwgll(z, w, lx)
    return z, w


def dgll(z):
    lx = len(z)
    d = np.zeros(lx * lx)
    dt = np.zeros(lx * lx)
    lib().or_dgll(d, dt, np.ascontiguousarray(z, dtype=np.float64), lx)
    return d, dt


def read_nmsh(path):
    """Neko .nmsh reader restated for the tests (nmsh_file.f90:70-207 with
    the record layouts of format/nmsh.f90): int32 nelv, int32 gdim; nelv
    packed hex records (int32 el_idx, 8 x (int32 v_idx, 3 x f64)); int32
    nzones and 9-int32 zone records (e, f, p_e, p_f, glb_pt_ids(4), type);
    int32 ncurves.  Vertices are returned in Neko's symmetric order (file
    p1,p2,p4,p3,p5,p6,p8,p7 -> Neko 1..8, nmsh_file.f90:159); periodic zones
    (type 5) replace the ids of their facet points by glb_pt_ids, on the
    point object, i.e. in every element holding that point
    (mesh_apply_periodic_facet, mesh.f90:1837-1866).  Returns (vxyz (nel,8,
    3), pid (nel,8), zones (nz,9), ncurves)."""
    b = open(path, "rb").read()
    nel, gdim = np.frombuffer(b, dtype="<i4", count=2)
    if gdim != 3:
        raise ValueError("only 3-D meshes")
    rec = np.dtype([("el", "<i4"), ("v", [("idx", "<i4"), ("xyz", "<f8", 3)], 8)])
    h = np.frombuffer(b, dtype=rec, count=int(nel), offset=8)
    off = 8 + int(nel) * rec.itemsize
    nz = int(np.frombuffer(b, dtype="<i4", count=1, offset=off)[0])
    zones = np.frombuffer(b, dtype="<i4", count=nz * 9, offset=off + 4).reshape(nz, 9)
    ncur = int(np.frombuffer(b, dtype="<i4", count=1, offset=off + 4 + 36 * nz)[0])
    order = [0, 1, 3, 2, 4, 5, 7, 6]
    vxyz = np.ascontiguousarray(h["v"]["xyz"][:, order, :], dtype=np.float64)
    pid = np.ascontiguousarray(h["v"]["idx"][:, order], dtype=np.int64)
    face_nodes = [[1, 5, 7, 3], [2, 6, 8, 4], [1, 2, 6, 5], [3, 4, 8, 7],
                  [1, 2, 4, 3], [5, 6, 8, 7]]
    loc = {int(g): e for e, g in enumerate(h["el"])}
    new_id = {}
    for z in zones:
        if z[8] != 5:
            continue
        e = loc[int(z[0])]
        for c in range(4):
            new_id[int(pid[e, face_nodes[z[1] - 1][c] - 1])] = int(z[4 + c])
    if new_id:
        pid = np.array([[new_id.get(int(q), int(q)) for q in row] for row in pid],
                       dtype=np.int64)
    return vxyz, pid, zones.copy(), ncur


def read_nmsh_curves(path):
    """The curve records of a .nmsh file (nmsh_file_read_curves,
    nmsh_file.f90:575-639; nmsh_curve_el_t, format/nmsh.f90:39-43, written
    through the MPI struct type of mpi_types.f90:225-249, i.e. packed: int32
    e, 5 x 12 f64 curve_data (column-major: 5 values per edge), 12 int32
    edge types; 532 bytes).  Returns (el (nc,) 0-based file-order element
    of each record, data (nc, 60), type (nc, 12))."""
    b = open(path, "rb").read()
    nel = int(np.frombuffer(b, dtype="<i4", count=1)[0])
    hx = np.dtype([("el", "<i4"), ("v", [("idx", "<i4"), ("xyz", "<f8", 3)], 8)])
    els = np.frombuffer(b, dtype=hx, count=nel, offset=8)["el"]
    off = 8 + nel * hx.itemsize
    nz = int(np.frombuffer(b, dtype="<i4", count=1, offset=off)[0])
    off += 4 + 36 * nz
    nc = int(np.frombuffer(b, dtype="<i4", count=1, offset=off)[0])
    rec = np.dtype([("e", "<i4"), ("cd", "<f8", 60), ("t", "<i4", 12)])
    assert rec.itemsize == 532
    r = np.frombuffer(b, dtype=rec, count=nc, offset=off + 4)
    loc = {int(g): e for e, g in enumerate(els)}
    el = np.array([loc[int(e)] for e in r["e"]], dtype=np.int32)
    return el, np.ascontiguousarray(r["cd"]), np.ascontiguousarray(r["t"], dtype=np.int32)


class OracleCase:
    """A whole Neko-style problem on a (possibly perturbed/periodic) box."""

    def __init__(self, nelx, nely, nelz, lx, amp=0.0, seed=0,
                 periodic=(False, False, False), h1=1.0, h2=0.0):
        L = lib()
        self.nelx, self.nely, self.nelz = nelx, nely, nelz
        self.amp, self.seed, self.periodic = amp, seed, tuple(periodic)
        nel = nelx * nely * nelz
        vxyz = np.zeros(nel * 24)
        pid = np.zeros(nel * 8, dtype=np.int64)
        L.or_box_mesh(nelx, nely, nelz, int(periodic[0]), int(periodic[1]),
                      int(periodic[2]), float(amp), int(seed), vxyz, pid)
        self._setup(vxyz, pid, lx, h1, h2)

    @classmethod
    def from_mesh(cls, vxyz, pid, lx, h1=1.0, h2=0.0, curves=None):
        """The same problem on a general conforming hex mesh (vertices in
        Neko's symmetric order, point ids), Dirichlet on every facet without
        a neighbour; curves = (el, data, type) as read_nmsh_curves returns
        them, or None for straight edges."""
        c = cls.__new__(cls)
        c.nelx = c.nely = c.nelz = None
        c.amp, c.seed, c.periodic = None, None, None
        c.curves = curves
        c._setup(np.ascontiguousarray(vxyz, dtype=np.float64).reshape(-1),
                 np.ascontiguousarray(pid, dtype=np.int64).reshape(-1), lx,
                 h1, h2)
        return c

    def same_mesh(self, lx, h1=1.0, h2=0.0):
        """The problem on this case's mesh at another order (multigrid
        levels, the Schwarz extended space)."""
        if self.nelx is not None:
            return OracleCase(self.nelx, self.nely, self.nelz, lx, amp=self.amp,
                              seed=self.seed, periodic=self.periodic, h1=h1, h2=h2)
        return OracleCase.from_mesh(self.vxyz, self.pid, lx, h1=h1, h2=h2,
                                    curves=getattr(self, "curves", None))

    def _setup(self, vxyz, pid, lx, h1, h2):
        L = lib()
        self.lx = lx
        nel = len(pid) // 8
        self.nel = nel
        self.n = n = nel * lx ** 3
        self.vxyz = vxyz
        self.pid = pid
        self.fneigh = np.zeros(nel * 6, dtype=np.int32)
        L.or_facet_neigh(nel, self.pid, self.fneigh)
        self.dfrmd = np.zeros(nel, dtype=np.int32)
        L.or_dfrmd_el(nel, self.vxyz, self.dfrmd)
        self.z, self.w = zwgll(lx)
        self.d, self.dt = dgll(self.z)
        w = self.w
        self.w3 = np.einsum("k,j,i->kji", w, w, w).reshape(-1).copy()
        self.dof = np.zeros(n, dtype=np.int64)
        L.or_dof_ids(nel, lx, self.pid, self.dof)
        self.x = np.zeros(n)
        self.y = np.zeros(n)
        self.zc = np.zeros(n)
        curves = getattr(self, "curves", None)
        if curves is not None and len(curves[0]):
            cel = np.ascontiguousarray(curves[0], dtype=np.int32)
            cdata = np.ascontiguousarray(curves[1], dtype=np.float64).reshape(-1)
            ctype = np.ascontiguousarray(curves[2], dtype=np.int32).reshape(-1)
            st = C.c_int()
            L.or_xyz_curved(nel, lx, self.z, self.vxyz, len(cel), cel, cdata, ctype,
                            self.x, self.y, self.zc, C.byref(st))
            if st.value:
                raise ValueError("curve radius too small for an arc edge")
        else:
            L.or_xyz(nel, lx, self.z, self.vxyz, self.x, self.y, self.zc)
        # gs maps
        cap = n
        self.dg = np.zeros(cap, dtype=np.int32)
        self.gd = np.zeros(cap, dtype=np.int32)
        bl = np.zeros(cap, dtype=np.int32)
        bo = np.zeros(cap, dtype=np.int32)
        nl, fo, nb = C.c_int(), C.c_int(), C.c_int()
        L.or_gs_map(nel, lx, self.dof, self.fneigh, self.dg, self.gd,
                    C.byref(nl), C.byref(fo), bl, bo, C.byref(nb))
        self.nlocal, self.facet_offset, self.nblks = nl.value, fo.value, nb.value
        self.dg = self.dg[: self.nlocal].copy()
        self.gd = self.gd[: self.nlocal].copy()
        self.blk_len = bl[: self.nblks].copy()
        self.blk_off = bo[: self.nblks].copy()
        self.gs_buf = np.zeros(max(self.nlocal, 1))
        # coef
        self.g = [np.zeros(n) for _ in range(6)]
        self.jac = np.zeros(n)
        self.B = np.zeros(n)
        L.or_coef(nel, lx, self.x, self.y, self.zc, self.d, self.w3, *self.g,
                  self.jac, self.B)
        self.mult = np.ones(n)
        self.gs(self.mult)
        self.mult = 1.0 / self.mult
        self.volume = float(np.sum(self.B))
        self.h1 = np.full(n, float(h1))
        self.h2 = np.full(n, float(h2))
        self.ifh2 = h2 != 0.0
        m = L.or_mask(nel, lx, self.fneigh, None)
        self.msk = np.zeros(max(m, 1), dtype=np.int32)
        L.or_mask(nel, lx, self.fneigh, self.msk.ctypes.data_as(C.c_void_p))
        self.nmsk = m
        self._prob = None

    # ---- operations -----------------------------------------------------
    def gs(self, u):
        lib().or_gs_op_add(u, self.n, self.nlocal, self.facet_offset, self.dg,
                           self.gd, self.nblks, self.blk_len, self.blk_off,
                           self.gs_buf)
        return u

    def ax_local(self, u, h1=None):
        w = np.zeros(self.n)
        lib().or_ax_helm(w, u, self.d, self.dt,
                         self.h1 if h1 is None else h1, *self.g[:3],
                         *self.g[3:], self.nel, self.lx, self.h2, self.B,
                         int(self.ifh2))
        return w

    def mask_apply(self, u):
        lib().or_mask_apply(u, self.msk, self.nmsk)
        return u

    def glsc3(self, a, b, c):
        return lib().or_glsc3(a, b, c, len(a))

    def jacobi(self):
        dd = np.zeros(self.n)
        lib().or_jacobi_diag(dd, self.dt, *self.g, self.dfrmd, self.nel,
                             self.lx)
        dd = dd * self.h1
        if self.ifh2:
            dd = dd + self.h2 * self.B
        self.gs(dd)
        return 1.0 / dd

    def problem(self):
        if self._prob is None:
            p = OrProblem()
            p.nel, p.lx = self.nel, self.lx
            p.d = self.d.ctypes.data
            p.dt = self.dt.ctypes.data
            p.h1 = self.h1.ctypes.data
            p.h2 = self.h2.ctypes.data
            p.B = self.B.ctypes.data
            p.ifh2 = int(self.ifh2)
            for name, arr in zip(("g11", "g22", "g33", "g12", "g13", "g23"),
                                 self.g):
                setattr(p, name, arr.ctypes.data)
            p.mult = self.mult.ctypes.data
            p.gs_m, p.gs_o, p.gs_nb = self.nlocal, self.facet_offset, self.nblks
            p.dg = self.dg.ctypes.data
            p.gd = self.gd.ctypes.data
            p.blk_len = self.blk_len.ctypes.data
            p.blk_off = self.blk_off.ctypes.data
            p.gs_buf = self.gs_buf.ctypes.data
            p.msk = self.msk.ctypes.data
            p.nmsk = self.nmsk
            p.volume = self.volume
            self._prob = p
        return self._prob

    def ax_full(self, u):
        w = np.zeros(self.n)
        lib().or_ax_full(C.byref(self.problem()), w,
                         np.ascontiguousarray(u, dtype=np.float64))
        return w

    def rhs_normal(self, seed):
        """i.i.d. N(0,1) per local dof, made consistent by one gs sum, then
        masked (the synthetic RHS of the large Poisson workload)."""
        L = lib()
        f = np.array([L.or_rng_normal(seed, i) for i in range(self.n)])
        self.gs(f)
        self.mask_apply(f)
        return f

    def rhs_sine(self):
        """b = mask(gs(B * 3 pi^2 sin(pi x) sin(pi y) sin(pi z)))."""
        pi = np.pi
        f = 3 * pi * pi * np.sin(pi * self.x) * np.sin(pi * self.y) * \
            np.sin(pi * self.zc)
        f = self.B * f
        self.gs(f)
        self.mask_apply(f)
        return f

    def rhs_poisson_example(self):
        """set_f of examples/poisson/setup.f90: f = 500 exp(-|X - (4,4,4)|^2
        / 2) per local point, gs sum, times mult; then the Dirichlet mask
        (driver.f90: bclst%apply(f))."""
        r2 = (self.x - 4.0) ** 2 + (self.y - 4.0) ** 2 + (self.zc - 4.0) ** 2
        f = 500.0 * np.exp(-r2 / 2.0)
        self.gs(f)
        f = f * self.mult
        self.mask_apply(f)
        return f

    def cofactors(self):
        """coef's drdx, drdy, drdz, dsdx, .., dtdz as a (9, n) array."""
        if getattr(self, "_cof", None) is None:
            cof = np.zeros(9 * self.n)
            lib().or_cofactors(self.nel, self.lx, self.x, self.y, self.zc, self.d, cof)
            self._cof = cof
        return self._cof.reshape(9, self.n)

    def opgrad(self, u):
        """opgrad (cpu_opgrad_lx): the weak gradient w3 * C^T grad_r u."""
        ux, uy, uz = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        lib().or_opgrad(self.nel, self.lx, np.ascontiguousarray(u, dtype=np.float64),
                        self.d, self.w3, np.ascontiguousarray(self.cofactors()).reshape(-1),
                        ux, uy, uz)
        return ux, uy, uz

    def dudxyz(self, u, c):
        """du/dx_c (cpu_dudxyz_lx with coef's dr_c, ds_c, dt_c, jacinv)."""
        cof = self.cofactors()
        du = np.zeros(self.n)
        lib().or_dudxyz(self.nel, self.lx, np.ascontiguousarray(u, dtype=np.float64),
                        np.ascontiguousarray(cof[c]), np.ascontiguousarray(cof[3 + c]),
                        np.ascontiguousarray(cof[6 + c]), self.jac, self.d, du)
        return du

    def cdtp(self, x, c):
        """cdtp(x, dr_c, ds_c, dt_c) (cpu_cdtp_lx)."""
        cof = self.cofactors()
        out = np.zeros(self.n)
        lib().or_cdtp(self.nel, self.lx, np.ascontiguousarray(x, dtype=np.float64),
                      np.ascontiguousarray(cof[c]), np.ascontiguousarray(cof[3 + c]),
                      np.ascontiguousarray(cof[6 + c]), self.d, self.w3, out)
        return out

    def binv(self):
        b = self.B.copy()
        self.gs(b)
        return 1.0 / b

    def curl(self, u1, u2, u3):
        """opr_cpu_curl (opr_cpu.f90:126-180): w1 = dw/dy - dv/dz, ..., then
        w *= B, gs, w *= Binv (no cyclic rotation)."""
        w1 = self.dudxyz(u3, 1) - self.dudxyz(u2, 2)
        w2 = self.dudxyz(u1, 2) - self.dudxyz(u3, 0)
        w3 = self.dudxyz(u2, 0) - self.dudxyz(u1, 1)
        bi = self.binv()
        out = []
        for w in (w1, w2, w3):
            w = w * self.B
            self.gs(w)
            out.append(w * bi)
        return out

    def prs_res(self, p, ue, f, mu, rho):
        """pnpn_prs_res_cpu_compute (pnpn_res_cpu.f90:31-180) with h1 = 1/rho,
        h2 = 0 and no surface terms (no velocity-prescribed or symmetry
        facets carrying a normal flux)."""
        ta = self.curl(*ue)
        wa = self.curl(*ta)
        bi = self.binv()
        ta = [f[c] / rho - ((wa[c] * (mu / rho)) * self.B) for c in range(3)]
        for t in ta:
            self.gs(t)
        ta = [t * bi for t in ta]
        wa = [self.cdtp(ta[c], c) for c in range(3)]
        pr = self.ax_local_coeffs(p, 1.0 / rho, 0.0)
        return (-pr) + wa[0] + wa[1] + wa[2]

    def cyclic_setup(self, facets):
        """coef_cyclic (src/sem/coef.f90:1380-1430): for each cyclic facet in
        list order and its points in Neko's i, j, k loop order, the unit
        normal n (coef_get_normal) gives R11 = n_x/|n_xy| sign(d), R12 =
        n_y/|n_xy| sign(d), d = y n_x - x n_y (Fortran SIGN(1, d): the sign
        bit of d).  Returns (cyc_msk 1-based, R11, R12)."""
        pts = []
        self.facet_normals(facets, points=pts)
        msk, r11, r12 = [], [], []
        eps = np.finfo(np.float64).eps
        for loc, nx, ny, _ in pts:
            ln = np.sqrt(nx * nx + ny * ny)
            if not ln > eps:
                raise ValueError("normal without an x-y part")
            d = self.y[loc] * nx - self.x[loc] * ny
            sg = np.copysign(1.0, d)
            msk.append(loc + 1)
            r11.append(nx / ln * sg)
            r12.append(ny / ln * sg)
        return np.array(msk, dtype=np.int64), np.array(r11), np.array(r12)

    @staticmethod
    def rotate_cyc(vx, vy, idir, setup):
        """opr_cpu_rotate_cyc_r1 (src/math/bcknd/cpu/opr_cpu.f90:333-357),
        in place, entries in order."""
        msk, r11, r12 = setup
        for i in range(len(msk)):
            j = msk[i] - 1
            if idir == 1:
                vnor = vx[j] * r11[i] + vy[j] * r12[i]
                vtan = -vx[j] * r12[i] + vy[j] * r11[i]
            else:
                vnor = vx[j] * r11[i] - vy[j] * r12[i]
                vtan = vx[j] * r12[i] + vy[j] * r11[i]
            vx[j] = vnor
            vy[j] = vtan

    def facet_normals(self, facets, points=None):
        """facet_normal_t over a list of (element, facet 1..6) pairs
        (src/bc/facet_normal.f90:233-326): the unique local points of the
        facets in first-seen order and their area-weighted normals summed over
        the facets in list order, the normal and area of each facet point as
        coef_generate_area_and_normal (src/sem/coef.f90:1195-1370) makes them:
        cross products of the tangent derivatives (D applied along the face's
        two directions), signs -,+ / +,- / -,+ for facets 1,2 / 3,4 / 5,6,
        area = |cross| w_a w_b, normal = cross / |cross| above machine
        epsilon."""
        lx = self.lx
        D = self.d.reshape(lx, lx, order="F")  # D[i, j] = d[i + lx j]
        X = [c.reshape(self.nel, lx, lx, lx) for c in (self.x, self.y, self.zc)]  # [e, k, j, i]
        w = self.w
        eps = np.finfo(np.float64).eps
        slot = {}
        pts, nrm = [], []
        for e, f in facets:
            e = int(e)
            f = int(f)
            xe = [c[e] for c in X]
            # derivatives along r (i), s (j), t (k) of x, y, z: arrays [k, j, i]
            dr = [np.einsum("im,kjm->kji", D, c) for c in xe]
            ds = [np.einsum("jm,kmi->kji", D, c) for c in xe]
            dt = [np.einsum("km,mji->kji", D, c) for c in xe]
            if f in (1, 2):
                t1, t2, sgn = ds, dt, (-1.0 if f == 1 else 1.0)
                sel = (slice(None), slice(None), 0 if f == 1 else lx - 1)
            elif f in (3, 4):
                t1, t2, sgn = dr, dt, (1.0 if f == 3 else -1.0)
                sel = (slice(None), 0 if f == 3 else lx - 1, slice(None))
            else:
                t1, t2, sgn = dr, ds, (-1.0 if f == 5 else 1.0)
                sel = (0 if f == 5 else lx - 1, slice(None), slice(None))
            a = t1[1] * t2[2] - t1[2] * t2[1]
            b = t1[2] * t2[0] - t1[0] * t2[2]
            c = t1[0] * t2[1] - t1[1] * t2[0]
            a, b, c = a[sel], b[sel], c[sel]  # 2-D over the face's (slow, fast) axes
            dot = a * a + b * b + c * c
            wgt = np.outer(w, w)  # weight of (slow, fast) face coordinates
            area = np.sqrt(dot) * wgt
            nx, ny, nz = sgn * a, sgn * b, sgn * c
            ln = np.sqrt(nx ** 2 + ny ** 2 + nz ** 2)
            big = ln > eps
            nx = np.where(big, nx / np.where(big, ln, 1.0), nx)
            ny = np.where(big, ny / np.where(big, ln, 1.0), ny)
            nz = np.where(big, nz / np.where(big, ln, 1.0), nz)
            kk, jj, ii = np.meshgrid(np.arange(lx), np.arange(lx), np.arange(lx), indexing="ij")
            loc = (e * lx ** 3 + ii + lx * (jj + lx * kk))[sel]
            if points is not None:
                for q in range(lx * lx):
                    s_, f_ = divmod(q, lx)
                    points.append((int(loc[s_, f_]), nx[s_, f_], ny[s_, f_], nz[s_, f_]))
                continue
            for q in range(lx * lx):
                s_, f_ = divmod(q, lx)
                k = int(loc[s_, f_])
                if k not in slot:
                    slot[k] = len(pts)
                    pts.append(k)
                    nrm.append([0.0, 0.0, 0.0])
                v = nrm[slot[k]]
                v[0] = v[0] + nx[s_, f_] * area[s_, f_]
                v[1] = v[1] + ny[s_, f_] * area[s_, f_]
                v[2] = v[2] + nz[s_, f_] * area[s_, f_]
        nrm = np.array(nrm, dtype=np.float64).reshape(-1, 3)
        return np.array(pts, dtype=np.int64), nrm[:, 0].copy(), nrm[:, 1].copy(), nrm[:, 2].copy()

    def prs_res_surf(self, p, ue, u, f, mu, rho, bd, dt, prs=None, sym=None):
        """pnpn_prs_res_cpu_compute (pnpn_res_cpu.f90:31-180) with the surface
        terms: p_res -= (bd/dt) (u . n A) on the velocity-prescribed facets'
        points and -= (ta . n A) on the symmetry facets' points, n A from
        facet_normals, ta = Binv gs(f/rho - (mu/rho) B curl curl u_e)
        (facet_normal_apply_surfvec, src/bc/facet_normal.f90:146-171)."""
        ta = self.curl(*ue)
        wa = self.curl(*ta)
        bi = self.binv()
        ta = [f[c] / rho - ((wa[c] * (mu / rho)) * self.B) for c in range(3)]
        for t in ta:
            self.gs(t)
        ta = [t * bi for t in ta]
        wa = [self.cdtp(ta[c], c) for c in range(3)]
        pr = self.ax_local_coeffs(p, 1.0 / rho, 0.0)
        out = (-pr) + wa[0] + wa[1] + wa[2]
        dtbd = bd / dt
        tp = [np.zeros(self.n) for _ in range(3)]
        if prs is not None:
            pts, nx, ny, nz = prs
            tp[0][pts] = u[0][pts] * nx
            tp[1][pts] = u[1][pts] * ny
            tp[2][pts] = u[2][pts] * nz
        ts = [np.zeros(self.n) for _ in range(3)]
        if sym is not None:
            pts, nx, ny, nz = sym
            ts[0][pts] = ta[0][pts] * nx
            ts[1][pts] = ta[1][pts] * ny
            ts[2][pts] = ta[2][pts] * nz
        return out - (dtbd * (tp[0] + tp[1] + tp[2])) - (ts[0] + ts[1] + ts[2])

    def box_facets(self, dims, side):
        """(element, facet) pairs of one side of a box mesh built by
        OracleCase (element e = ex + nx (ey + ny ez)); side = facet number
        1..6 (x-, x+, y-, y+, z-, z+)."""
        nx, ny, nz = dims
        out = []
        for ez in range(nz):
            for ey in range(ny):
                for ex in range(nx):
                    on = {1: ex == 0, 2: ex == nx - 1, 3: ey == 0, 4: ey == ny - 1,
                          5: ez == 0, 6: ez == nz - 1}[side]
                    if on:
                        out.append((ex + nx * (ey + ny * ez), side))
        return out

    def ax_local_coeffs(self, u, h1, h2):
        """ax_helm with constant h1, h2 (h2 != 0 adds h2 B u), no gs."""
        w = np.zeros(self.n)
        h1a = np.full(self.n, float(h1))
        h2a = np.full(self.n, float(h2))
        lib().or_ax_helm(w, np.ascontiguousarray(u, dtype=np.float64), self.d, self.dt, h1a,
                         *self.g[:3], *self.g[3:], self.nel, self.lx, h2a, self.B,
                         int(h2 != 0.0))
        return w

    def vel_res(self, u, v, w, p, f, mu, rho, bd, dt):
        """pnpn_vel_res_cpu_compute (pnpn_res_cpu.f90:182-241): h1 = mu,
        h2 = rho bd / dt; res_c = -Ax(u_c) - opgrad_c(p) + f_c (local)."""
        h2 = rho * bd / dt
        g = self.opgrad(p)
        return [(-self.ax_local_coeffs(q, mu, h2)) - g[c] + f[c]
                for c, q in enumerate((u, v, w))]

    def res_norm(self, r):
        return float(np.sqrt(self.glsc3(r, self.mult, r)) /
                     np.sqrt(self.volume))

    def cg(self, f, abs_tol, max_iter=10000, pc=None, history=False):
        """Jacobi-PCG (cg.f90); history=True also returns rnorm after
        every iteration."""
        x = np.zeros(self.n)
        if pc is None:
            pc = self.jacobi()
        self._pc = pc
        hist = np.zeros(max_iter) if history else None
        if history:
            lib().or_cg_history(hist.ctypes.data_as(C.c_void_p), max_iter)
        try:
            mon = lib().or_cg_solve(C.byref(self.problem()),
                                    pc.ctypes.data_as(C.c_void_p), x,
                                    np.ascontiguousarray(f), max_iter,
                                    float(abs_tol))
        finally:
            if history:
                lib().or_cg_history(None, 0)
        if history:
            return x, mon, hist[: mon.iter].copy()
        return x, mon

    def gmres(self, f, abs_tol, max_iter=10000, lgmres=30, pc=None):
        x = np.zeros(self.n)
        pcp = None if pc is None else pc.ctypes.data_as(C.c_void_p)
        mon = lib().or_gmres_solve(C.byref(self.problem()), pcp, x,
                                   np.ascontiguousarray(f), max_iter,
                                   float(abs_tol), lgmres)
        return x, mon


# --------------------------------------------------------------------------
# hsmg (pc_hsmg.f90) on top of OracleCase levels
# --------------------------------------------------------------------------
class OrGs(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("o", C.c_int), ("nb", C.c_int),
                ("dg", C.c_void_p), ("gd", C.c_void_p), ("b", C.c_void_p),
                ("bo", C.c_void_p), ("buf", C.c_void_p)]


class OrSchwarz(C.Structure):
    _fields_ = [("lx", C.c_int), ("nel", C.c_int), ("gs", OrGs),
                ("gs_ext", OrGs), ("msk", C.c_void_p), ("nmsk", C.c_int),
                ("s", C.c_void_p), ("d", C.c_void_p), ("wt", C.c_void_p),
                ("work1", C.c_void_p), ("work2", C.c_void_p)]


class OrHsmg(C.Structure):
    _fields_ = [("nel", C.c_int), ("lf", C.c_int), ("lm", C.c_int),
                ("lc", C.c_int), ("nf", C.c_int64), ("nm", C.c_int64),
                ("nc", C.c_int64), ("sw_fine", OrSchwarz), ("sw_mid", OrSchwarz),
                ("gs_crs", OrGs), ("msk_c", C.c_void_p), ("nmsk_c", C.c_int),
                ("mult_f", C.c_void_p), ("mult_m", C.c_void_p),
                ("jfm", C.c_void_p), ("jfm_t", C.c_void_p), ("jmc", C.c_void_p),
                ("jmc_t", C.c_void_p), ("crs", OrProblem), ("pc_crs", C.c_void_p),
                ("niter", C.c_int), ("crs_tol", C.c_double), ("r", C.c_void_p),
                ("w", C.c_void_p), ("e_mid", C.c_void_p), ("e_crs", C.c_void_p),
                ("last_crs_iter", C.c_int)]


def _gs_struct(c):
    g = OrGs()
    g.n, g.m, g.o, g.nb = c.n, c.nlocal, c.facet_offset, c.nblks
    g.dg, g.gd = c.dg.ctypes.data, c.gd.ctypes.data
    g.b, g.bo = c.blk_len.ctypes.data, c.blk_off.ctypes.data
    g.buf = c.gs_buf.ctypes.data
    return g


def _bind_hsmg(L):
    if getattr(L, "_hsmg_bound", False):
        return
    L.or_setup_intp.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
    L.or_fdm_setup.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip,
                               C.POINTER(OrGs), _dp, _dp]
    L.or_schwarz_setup_wt.argtypes = [C.c_int, C.c_int, C.POINTER(OrGs),
                                      C.POINTER(OrGs), _dp]
    L.or_schwarz_compute.argtypes = [C.POINTER(OrSchwarz), _dp, _dp]
    L.or_hsmg_solve.argtypes = [C.POINTER(OrHsmg), _dp, _dp]
    L.or_gmres_hsmg_solve.restype = OrMonitor
    L.or_gmres_hsmg_solve.argtypes = [C.POINTER(OrProblem), C.POINTER(OrHsmg),
                                      _dp, _dp, C.c_int, C.c_double, C.c_int]
    L.or_tnsr3d.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int]
    L.or_semhat.argtypes = [C.c_int, _dp, _dp]
    L.or_fdm_1d.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double,
                            C.c_double, C.c_double, C.c_int]
    L.or_plane_space.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                 C.c_int]
    L._hsmg_bound = True


def semhat(n):
    """1-D reference stiffness ah ((n+1)^2, column-major) and GLL mass bh
    (fast3d.f90 semhat)."""
    L = lib()
    _bind_hsmg(L)
    ah = np.zeros((n + 1) * (n + 1))
    bh = np.zeros(n + 1)
    L.or_semhat(n, ah, bh)
    return ah.reshape(n + 1, n + 1).T.copy(), bh


def fdm_1d(lbc, rbc, ll, lm, lr, n):
    """Pencil (A, B) of the extended 1-D element (fdm.f90:478-560) and the
    FDM eigenvectors S (boundary rows zeroed) and eigenvalues."""
    L = lib()
    _bind_hsmg(L)
    nl = n + 3
    a = np.zeros(nl * nl)
    b = np.zeros(nl * nl)
    s = np.zeros(2 * nl * nl)
    lam = np.zeros(nl)
    L.or_fdm_1d(a, b, s, lam, lbc, rbc, ll, lm, lr, n)
    f = lambda m: m.reshape(nl, nl).T.copy()
    return f(a), f(b), f(s[: nl * nl]), f(s[nl * nl:]), lam


def plane_space(case):
    """Per-element lengths between the first and last GLL planes in r, s, t
    (fdm.f90 plane_space)."""
    L = lib()
    _bind_hsmg(L)
    out = [np.zeros(case.nel) for _ in range(3)]
    L.or_plane_space(*out, case.w, case.x, case.y, case.zc, case.lx, case.nel)
    return out


def tnsr3d(u, nu, nv, A, Bt, Ct, nel):
    """v = (C (x) B (x) A) u per element (tnsr3d_el_nvnu); matrices
    column-major numpy arrays (A: nv x nu, Bt, Ct: nu x nv)."""
    L = lib()
    _bind_hsmg(L)
    v = np.zeros(nel * nv ** 3)
    cm = lambda m: np.ascontiguousarray(np.asarray(m).T).ravel()
    L.or_tnsr3d(v, nv, np.ascontiguousarray(u, dtype=np.float64), nu, cm(A),
                cm(Bt), cm(Ct), nel)
    return v


def setup_intp(z_to, z_from):
    L = lib()
    _bind_hsmg(L)
    nt, nf = len(z_to), len(z_from)
    jh = np.zeros(nt * nf)
    jht = np.zeros(nt * nf)
    L.or_setup_intp(jh, jht, np.ascontiguousarray(z_to, dtype=np.float64),
                    np.ascontiguousarray(z_from, dtype=np.float64), nt, nf, 0)
    return jh, jht


class OracleHsmg:
    """hsmg with levels lx -> lx_mid -> 2 on the mesh of `case`
    (pc_hsmg.f90:147-378).  facet_bc: facet_type of boundary facets
    (1: Dirichlet as set for pressure, 2: as set for velocity)."""

    def __init__(self, case, facet_bc=2, niter=10, crs_tol=1e-9):
        L = lib()
        _bind_hsmg(L)
        self.case = case
        lf = case.lx
        if lf < 5:
            lm = max(lf - 1, 3)
        else:
            lm = 4
        lc = 2
        self.mid = case.same_mesh(lm)
        self.crs = case.same_mesh(lc, h1=float(case.h1[0]), h2=0.0)
        self.crs.dfrmd[:] = 1  # hsmg_init: msh%all_deformed()
        self.ext_f = case.same_mesh(lf + 2)
        self.ext_m = case.same_mesh(lm + 2)
        nel = case.nel
        self.lf, self.lm, self.lc = lf, lm, lc
        ft = np.where(case.fneigh != 0, 0, facet_bc).astype(np.int32)
        self.facet_type = ft
        self._keep = []

        def schwarz(c, ext):
            nl = c.lx + 2
            s = np.zeros(nel * nl * nl * 6)
            d = np.zeros(nel * nl ** 3)
            g = _gs_struct(c)
            ge = _gs_struct(ext)
            L.or_fdm_setup(nel, c.lx, c.x, c.y, c.zc, c.w, ft, C.byref(g), s, d)
            wt = np.zeros(nel * c.lx * c.lx * 12)
            L.or_schwarz_setup_wt(nel, c.lx, C.byref(ge), C.byref(g), wt)
            w1 = np.zeros(nel * nl ** 3)
            w2 = np.zeros(nel * nl ** 3)
            sw = OrSchwarz()
            sw.lx, sw.nel, sw.gs, sw.gs_ext = c.lx, nel, g, ge
            sw.msk, sw.nmsk = c.msk.ctypes.data, c.nmsk
            sw.s, sw.d, sw.wt = s.ctypes.data, d.ctypes.data, wt.ctypes.data
            sw.work1, sw.work2 = w1.ctypes.data, w2.ctypes.data
            self._keep += [s, d, wt, w1, w2]
            return sw, s, d, wt

        self.sw_f, self.s_f, self.d_f, self.wt_f = schwarz(case, self.ext_f)
        self.sw_m, self.s_m, self.d_m, self.wt_m = schwarz(self.mid, self.ext_m)
        self.jfm, self.jfm_t = setup_intp(case.z, self.mid.z)
        self.jmc, self.jmc_t = setup_intp(self.mid.z, self.crs.z)
        self.pc_crs = self.crs.jacobi()
        h = OrHsmg()
        h.nel, h.lf, h.lm, h.lc = nel, lf, lm, lc
        h.nf, h.nm, h.nc = case.n, self.mid.n, self.crs.n
        h.sw_fine, h.sw_mid = self.sw_f, self.sw_m
        h.gs_crs = _gs_struct(self.crs)
        h.msk_c, h.nmsk_c = self.crs.msk.ctypes.data, self.crs.nmsk
        h.mult_f, h.mult_m = case.mult.ctypes.data, self.mid.mult.ctypes.data
        h.jfm, h.jfm_t = self.jfm.ctypes.data, self.jfm_t.ctypes.data
        h.jmc, h.jmc_t = self.jmc.ctypes.data, self.jmc_t.ctypes.data
        h.crs = self.crs.problem()
        h.pc_crs = self.pc_crs.ctypes.data
        h.niter, h.crs_tol = niter, crs_tol
        self.r = np.zeros(case.n)
        self.w = np.zeros(case.n)
        self.e_mid = np.zeros(self.mid.n)
        self.e_crs = np.zeros(self.crs.n)
        h.r, h.w = self.r.ctypes.data, self.w.ctypes.data
        h.e_mid, h.e_crs = self.e_mid.ctypes.data, self.e_crs.ctypes.data
        self.h = h

    def apply(self, r):
        z = np.zeros(self.case.n)
        lib().or_hsmg_solve(C.byref(self.h), z, np.ascontiguousarray(r, dtype=np.float64))
        return z

    def schwarz_fine(self, r):
        e = np.zeros(self.case.n)
        rr = np.array(r, dtype=np.float64)
        lib().or_schwarz_compute(C.byref(self.sw_f), e, rr)
        return e

    def schwarz_mid(self, r):
        e = np.zeros(self.mid.n)
        rr = np.array(r, dtype=np.float64)
        lib().or_schwarz_compute(C.byref(self.sw_m), e, rr)
        return e

    def gmres(self, f, abs_tol, max_iter=10000, lgmres=30):
        x = np.zeros(self.case.n)
        mon = lib().or_gmres_hsmg_solve(C.byref(self.case.problem()), C.byref(self.h),
                                        x, np.ascontiguousarray(f), max_iter,
                                        float(abs_tol), lgmres)
        return x, mon


# --------------------------------------------------------------------------
# Chebyshev smoother and phmg-style V-cycle (neko_oracle_phmg.c)
# --------------------------------------------------------------------------
class OrCheby(C.Structure):
    _fields_ = [("pb", C.c_void_p), ("sw", C.c_void_p), ("pcd", C.c_void_p),
                ("tha", C.c_double), ("dlt", C.c_double),
                ("d", C.c_void_p), ("w", C.c_void_p), ("r", C.c_void_p)]


class OrPhmg(C.Structure):
    _fields_ = [("nel", C.c_int), ("lf", C.c_int), ("lm", C.c_int), ("lc", C.c_int),
                ("pb", C.c_void_p * 3), ("ch", OrCheby * 2),
                ("smooth_its", C.c_int),
                ("jfm", C.c_void_p), ("jfm_t", C.c_void_p),
                ("jmc", C.c_void_p), ("jmc_t", C.c_void_p),
                ("pc_crs", C.c_void_p), ("crs_niter", C.c_int),
                ("crs_tol", C.c_double),
                ("r", C.c_void_p * 3), ("z", C.c_void_p * 3), ("w", C.c_void_p * 3),
                ("amg", C.c_void_p)]


def _bind_phmg(L):
    if getattr(L, "_phmg_bound", False):
        return
    L.or_cheby_power.argtypes = [C.POINTER(OrCheby), C.c_uint64, C.c_int]
    L.or_cheby_solve.argtypes = [C.POINTER(OrCheby), _dp, _dp, C.c_int, C.c_int]
    L.or_phmg_solve.argtypes = [C.POINTER(OrPhmg), _dp, _dp]
    L.or_gmres_phmg_solve.restype = OrMonitor
    L.or_gmres_phmg_solve.argtypes = [C.POINTER(OrProblem), C.POINTER(OrPhmg),
                                      _dp, _dp, C.c_int, C.c_double, C.c_int]
    L._phmg_bound = True


class OracleCheby:
    """Chebyshev smoother (cheby_t) on `case`: accelerated by a Schwarz
    smoother (an OracleHsmg's fine level) or by Jacobi."""

    def __init__(self, case, schwarz=None, seed=3901, power_its=150):
        L = lib()
        _bind_hsmg(L)
        _bind_phmg(L)
        self.case = case
        n = case.n
        self.bufs = [np.zeros(n) for _ in range(3)]
        c = OrCheby()
        c.pb = C.cast(C.pointer(case.problem()), C.c_void_p)
        self._sw = schwarz
        if schwarz is not None:
            c.sw = C.cast(C.pointer(schwarz), C.c_void_p)
            self.pcd = None
        else:
            self.pcd = case.jacobi()
            c.pcd = self.pcd.ctypes.data
        c.d, c.w, c.r = (b.ctypes.data for b in self.bufs)
        self.c = c
        L.or_cheby_power(C.byref(c), int(seed), int(power_its))

    @property
    def tha(self):
        return self.c.tha

    @property
    def dlt(self):
        return self.c.dlt

    def solve(self, f, niter, x0=None):
        x = np.zeros(self.case.n) if x0 is None else np.array(x0, dtype=np.float64)
        lib().or_cheby_solve(C.byref(self.c), x, np.ascontiguousarray(f), int(niter),
                             1 if x0 is None else 0)
        return x


class OraclePhmg:
    """phmg-style V-cycle on `case`: levels lx -> lx_mid -> 2 (hsmg's),
    Chebyshev(smooth_its) smoothing accelerated by the levels' Schwarz
    smoothers, coarse Jacobi-CG (niter iterations).  The level operators
    follow phmg's set-up: h1 of the fine level, ifh2 of the fine level and
    coef's default h2 = 1 on the coarser levels (level_h2="default"), or
    the fine level's h2 (level_h2="fine")."""

    def __init__(self, case, facet_bc=2, smooth_its=4, seed=3901, power_its=150,
                 crs_niter=10, crs_tol=1e-9, level_h2="default", coarse_amg=None):
        L = lib()
        _bind_hsmg(L)
        _bind_phmg(L)
        self.case = case
        self.hs = OracleHsmg(case, facet_bc=facet_bc)   # Schwarz + transfers
        hs = self.hs
        h1c = float(case.h1[0])
        # level_h2 "default": coef's h2 = 1 (phmg.f90 copies h1 and ifh2
        # only); "fine": the fine level's h2 on every level
        if level_h2 == "fine":
            h2l = float(case.h2[0]) if case.ifh2 else 0.0
        else:
            h2l = 1.0 if case.ifh2 else 0.0
        self.mid = case.same_mesh(hs.lm, h1=h1c, h2=h2l)
        self.crs = case.same_mesh(hs.lc, h1=h1c, h2=h2l)
        self.pc_crs = self.crs.jacobi()
        self.ch = [OracleCheby(case, hs.sw_f, seed, power_its),
                   OracleCheby(self.mid, hs.sw_m, seed, power_its)]
        h = OrPhmg()
        h.nel, h.lf, h.lm, h.lc = case.nel, hs.lf, hs.lm, hs.lc
        cases = (case, self.mid, self.crs)
        for l, c in enumerate(cases):
            h.pb[l] = C.cast(C.pointer(c.problem()), C.c_void_p)
        h.ch[0], h.ch[1] = self.ch[0].c, self.ch[1].c
        h.smooth_its = smooth_its
        h.jfm, h.jfm_t = hs.jfm.ctypes.data, hs.jfm_t.ctypes.data
        h.jmc, h.jmc_t = hs.jmc.ctypes.data, hs.jmc_t.ctypes.data
        h.pc_crs = self.pc_crs.ctypes.data
        h.crs_niter, h.crs_tol = crs_niter, crs_tol
        self.work = [[np.zeros(c.n) for c in cases] for _ in range(3)]
        for l in range(3):
            h.r[l] = self.work[0][l].ctypes.data
            h.z[l] = self.work[1][l].ctypes.data
            h.w[l] = self.work[2][l].ctypes.data
        # coarse solve: phmg's tree-AMG (levels, iterations, Chebyshev
        # degree; phmg.f90:150-157 defaults (3, 1, 4)) or hsmg's CG
        self.amg = None
        if coarse_amg is not None:
            self.amg = OracleTamg(self.crs, *coarse_amg)
            h.amg = self.amg.h
        self.h = h

    def apply(self, r):
        z = np.zeros(self.case.n)
        lib().or_phmg_solve(C.byref(self.h), z, np.ascontiguousarray(r, dtype=np.float64))
        return z

    def gmres(self, f, abs_tol, max_iter=10000, lgmres=30):
        x = np.zeros(self.case.n)
        mon = lib().or_gmres_phmg_solve(C.byref(self.case.problem()), C.byref(self.h),
                                        x, np.ascontiguousarray(f), max_iter,
                                        float(abs_tol), lgmres)
        return x, mon


# --------------------------------------------------------------------------
# tree-AMG (tamg_solver_t, src/multigrid/tree_amg*.f90; neko_oracle_tamg.c)
# --------------------------------------------------------------------------
def _bind_tamg(L):
    if getattr(L, "_tamg_bound", False):
        return
    L.or_tamg_new.restype = C.c_void_p
    L.or_tamg_new.argtypes = [C.POINTER(OrProblem), _ip, C.c_int, C.c_int, C.c_int,
                              C.c_int]
    L.or_tamg_delete.argtypes = [C.c_void_p]
    L.or_tamg_levels.restype = C.c_int
    L.or_tamg_levels.argtypes = [C.c_void_p, _lp]
    L.or_tamg_map.argtypes = [C.c_void_p, C.c_int, _ip]
    L.or_tamg_eigs.argtypes = [C.c_void_p, _dp, _dp]
    L.or_tamg_solve.argtypes = [C.c_void_p, _dp, _dp]
    L.or_tamg_matvec.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
    L.or_tamg_cheby.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int]
    L._tamg_bound = True


class OracleTamg:
    """tree-AMG on the problem of `case` (its mesh's facet graph): levels,
    V-cycles per solve, Chebyshev degree, power iterations."""

    def __init__(self, case, nlvls=3, max_iter=1, cheby_degree=4, power_its=250):
        L = lib()
        _bind_tamg(L)
        self.case = case
        self._fn = np.ascontiguousarray(case.fneigh, dtype=np.int32)
        self.h = L.or_tamg_new(C.byref(case.problem()), self._fn, int(nlvls),
                               int(max_iter), int(cheby_degree), int(power_its))
        if not self.h:
            raise ValueError("or_tamg_new failed")
        nn = np.zeros(16, dtype=np.int64)
        self.nlvls = L.or_tamg_levels(self.h, nn)
        self.nn = nn[: self.nlvls + 1].copy()

    def __del__(self):
        try:
            lib().or_tamg_delete(self.h)
        except Exception:
            pass

    def map(self, lvl):
        out = np.zeros(self.case.n, dtype=np.int32)
        lib().or_tamg_map(self.h, int(lvl), out)
        return out

    def eigs(self):
        t = np.zeros(self.nlvls)
        d = np.zeros(self.nlvls)
        lib().or_tamg_eigs(self.h, t, d)
        return t, d

    def matvec(self, v, lvl):
        out = np.zeros(int(self.nn[lvl]))
        lib().or_tamg_matvec(self.h, out, np.ascontiguousarray(v, dtype=np.float64), int(lvl))
        return out

    def cheby(self, f, lvl, x0=None):
        x = np.zeros(int(self.nn[lvl])) if x0 is None else np.array(x0, dtype=np.float64)
        lib().or_tamg_cheby(self.h, x, np.ascontiguousarray(f, dtype=np.float64), int(lvl),
                            1 if x0 is None else 0)
        return x

    def solve(self, r):
        z = np.zeros(self.case.n)
        lib().or_tamg_solve(self.h, z, np.ascontiguousarray(r, dtype=np.float64))
        return z


# --------------------------------------------------------------------------
# Solution-space projection (projection_t, src/common/projection.f90, CPU
# path), numpy over an OracleCase.  Parity unpinned by reference vectors
# (the reference ships none for it); pinned by the invariants in
# tests/test_oracle_proj.py (A-orthonormal basis, B = A X, W-orthogonal
# projected RHS, A xbar = b - b_proj, dependent vectors dropped, restart).
# --------------------------------------------------------------------------
class OracleProjection:
    def __init__(self, case, L, activ_step=5, tol=1e-7):
        self.case, self.L, self.activ, self.tol = case, L, activ_step, tol
        n = case.n
        self.m = 0
        self.xx = np.zeros((max(L, 1), n))
        self.bb = np.zeros((max(L, 1), n))
        self.xbar = np.zeros(n)
        self.proj_res = 0.0
        self.proj_m = 0

    def clear(self):
        # bcknd_clear (:855-875)
        self.m = 0
        self.proj_m = 0
        self.xx[:] = 0.0
        self.bb[:] = 0.0

    def pre(self, b, tstep, dt_last_change=-1):
        # pre_solving (:236-280), b projected in place
        if not (tstep > self.activ and self.L > 0):
            return
        if dt_last_change >= 0:
            if dt_last_change == 0:
                self.clear()
                return
            if not (dt_last_change > self.activ - 1):
                return
        self.project_on(b)

    def post(self, x, tstep, dt_last_change=-1):
        # post_solving (:282-301)
        if not (tstep > self.activ and self.L > 0):
            return
        if dt_last_change >= 0 and not (dt_last_change > self.activ - 1):
            return
        self.project_back(x)

    def project_on(self, b):
        # cpu_project_on (:476-551): two rounds of classical Gram-Schmidt
        m, w = self.m, self.case.mult
        if m <= 0:
            return
        self.proj_res = float(np.sqrt(np.dot(b * w, b) / self.case.volume))
        self.proj_m = m
        alpha = np.array([np.dot(self.xx[k] * w, b) for k in range(m)])
        xbar = alpha[0] * self.xx[0]
        b -= alpha[0] * self.bb[0]
        for k in range(1, m):
            xbar = xbar + alpha[k] * self.xx[k]
            b -= alpha[k] * self.bb[k]
        work = np.array([np.dot(self.xx[k] * w, b) for k in range(m)])
        for k in range(m):
            xbar = xbar + work[k] * self.xx[k]
            b -= work[k] * self.bb[k]
        self.xbar = xbar

    def project_back(self, x):
        # bcknd_project_back (:317-358), CPU branch
        if self.m > 0:
            x += self.xbar
        self.m = 1 if self.m == self.L else min(self.m + 1, self.L)
        m = self.m
        self.xx[m - 1] = x
        self.bb[m - 1] = self.case.ax_full(x)
        self.ortho()

    def ortho(self):
        # cpu_proj_ortho (:717-834)
        m, w = self.m, self.case.mult
        if m <= 0:
            return
        xm, bm = self.xx[m - 1], self.bb[m - 1]
        alpha = np.array([0.5 * (np.dot(self.xx[k] * w, bm) + np.dot(self.bb[k] * w, xm))
                          for k in range(m)])
        nrm = np.sqrt(alpha[m - 1])
        for k in range(m - 1):
            xm -= alpha[k] * self.xx[k]
            bm -= alpha[k] * self.bb[k]
        beta = np.array([0.5 * (np.dot(self.xx[k] * w, bm) + np.dot(self.bb[k] * w, xm))
                         for k in range(m - 1)])
        for k in range(m - 1):
            xm -= beta[k] * self.xx[k]
            bm -= beta[k] * self.bb[k]
        am = np.sqrt(np.dot(xm * w, bm))
        if am > self.tol * nrm:
            xm *= 1.0 / am
            bm *= 1.0 / am
        else:
            self.m = m - 1
