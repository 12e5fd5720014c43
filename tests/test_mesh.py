"""CPU tests of the general-mesh path: Neko .nmsh files (the reference's
own meshes, copied under tests/golden/nmsh) read by the product
(nm_nmsh_read) and by the oracle's reader; facet connectivity, gs maps and
mask of general conforming meshes (nm_mesh_*) against the oracle.

The oracle reader is pinned by the files themselves: every mesh lists its
boundary facets as zones (wall / labeled zones), so the facets that end up
without a neighbour must be exactly the non-periodic zone facets -- which
fixes the vertex swap, Neko's facet numbering, the periodic id replacement
and the connectivity at once.  The gs maps are pinned geometrically: two
local points share a gs id iff they are the same physical point (modulo the
period on periodic meshes)."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "nmsh")
MESHES = {  # fixture: (reference path, periodic box extent or None)
    "poisson_128.nmsh": ("examples/poisson/data/128.nmsh", None),
    "hemi.nmsh": ("examples/hemi/hemi.nmsh", None),
    "lid.nmsh": ("examples/lid/lid.nmsh", "auto"),
    "tgv_512.nmsh": ("examples/tgv/512.nmsh", 8.0),
}


def _path(name):
    return os.path.join(GOLD, name)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_fixture_is_the_reference_file(name):
    ref = os.path.join("/root/reference", MESHES[name][0])
    if not os.path.exists(ref):
        pytest.skip("reference checkout not present")
    a = hashlib.sha256(open(_path(name), "rb").read()).hexdigest()
    b = hashlib.sha256(open(ref, "rb").read()).hexdigest()
    assert a == b


@pytest.mark.parametrize("name", sorted(MESHES))
def test_oracle_reader_boundary_equals_zone_facets(name):
    vxyz, pid, zones, ncur = O.read_nmsh(_path(name))
    assert ncur == 0
    c = O.OracleCase.from_mesh(vxyz, pid, 3)
    bnd = {(e, f) for e in range(c.nel) for f in range(6)
           if c.fneigh[e * 6 + f] == 0}
    # zone element numbers are global ids; here they equal position + 1
    zone_facets = {(int(z[0]) - 1, int(z[1]) - 1) for z in zones if z[8] != 5}
    assert bnd == zone_facets
    if all(z[8] == 5 for z in zones):
        assert not bnd
    # right-handed elements after the vertex swap: positive Jacobian
    assert c.jac.min() > 0


def test_oracle_reader_box_geometry():
    vxyz, pid, zones, _ = O.read_nmsh(_path("poisson_128.nmsh"))
    assert vxyz.shape == (128, 8, 3)
    # lexicographic (symmetric) vertex order: vertex v at corner bits of v
    for v in range(8):
        cx, cy, cz = v & 1, (v >> 1) & 1, (v >> 2) & 1
        d = vxyz[:, v, :] - vxyz[:, 0, :]
        assert np.all((d[:, 0] > 0) == bool(cx))
        assert np.all((d[:, 1] > 0) == bool(cy))
        assert np.all((d[:, 2] > 0) == bool(cz))
    c = O.OracleCase.from_mesh(vxyz, pid, 4)
    assert abs(c.volume - 512.0) < 1e-10
    assert len(np.unique(pid)) == 9 * 5 * 5


def _period(name, vxyz):
    p = MESHES[name][1]
    if p is None:
        return None
    lo, hi = vxyz.reshape(-1, 3).min(0), vxyz.reshape(-1, 3).max(0)
    return hi - lo, lo


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("lx", [3, 5])
def test_oracle_gs_groups_are_coincident_points(name, lx):
    vxyz, pid, zones, _ = O.read_nmsh(_path(name))
    c = O.OracleCase.from_mesh(vxyz, pid, lx)
    X = np.stack([c.x, c.y, c.zc], axis=1)
    per = _period(name, vxyz)
    if per is not None:
        ext, lo = per
        # wrap every periodic direction (the coordinate extent of the
        # periodic meshes here is one period wherever zones connect)
        Xw = X.copy()
        for d in range(3):
            if ext[d] > 0:
                t = (Xw[:, d] - lo[d]) / ext[d]
                t = t - np.floor(t + 1e-9)
                Xw[:, d] = t * ext[d]
        X = Xw
    scale = np.abs(X).max() + 1.0
    key = np.round(X / (scale * 1e-9)).astype(np.int64)
    reg = c.gd - 1
    # same gs id -> same point
    order = np.argsort(c.dg, kind="stable")
    dg_s, pts = c.dg[order], X[reg[order]]
    starts = np.r_[0, np.nonzero(np.diff(dg_s))[0] + 1]
    for a, b in zip(starts, np.r_[starts[1:], len(dg_s)]):
        if b - a > 1:
            dev = np.abs(pts[a:b] - pts[a]).max()
            if per is not None:
                dev = min(dev, np.abs(np.abs(pts[a:b] - pts[a]) - ext).max())
            assert dev < 1e-9 * scale
    # same point -> same gs id (points never registered must be unique)
    gid = np.zeros(c.n, dtype=np.int64)
    gid[reg] = c.dg
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    if per is None:
        for grp in np.split(np.argsort(inv, kind="stable"),
                            np.nonzero(np.diff(np.sort(inv)))[0] + 1):
            if len(grp) > 1:
                ids = set(gid[grp].tolist())
                assert len(ids) == 1 and 0 not in ids


@pytest.mark.parametrize("name", sorted(MESHES))
def test_product_reader_matches_oracle_reader(name):
    from neko_amd.solve import read_nmsh
    v, p, el = read_nmsh(_path(name))
    vo, po, zones, _ = O.read_nmsh(_path(name))
    assert np.array_equal(v, vo)
    assert np.array_equal(p, po)
    assert np.array_equal(el, np.arange(1, len(el) + 1))


def test_product_reader_errors():
    from neko_amd._lib import NekoAmdError
    from neko_amd.solve import read_nmsh
    with pytest.raises(NekoAmdError, match="cannot open"):
        read_nmsh(os.path.join(GOLD, "missing.nmsh"))


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("lx", [2, 4, 8])
def test_product_mesh_gs_and_mask_match_oracle(name, lx):
    from neko_amd._lib import lib
    from neko_amd.solve import read_nmsh
    L = lib()
    v, p, _ = read_nmsh(_path(name))
    c = O.OracleCase.from_mesh(v, p, lx)
    nel = len(p)
    fn = np.zeros(nel * 6, dtype=np.int32)
    assert L.nm_mesh_facet_neigh(nel, p.ctypes.data, fn.ctypes.data) == 0
    assert np.array_equal(fn, c.fneigh)
    nl, fo, nb = C.c_int64(), C.c_int64(), C.c_int64()
    assert L.nm_mesh_gs_sizes(nel, lx, fn.ctypes.data, C.byref(nl), C.byref(fo)) == 0
    dg = np.zeros(nl.value, np.int32)
    gd = np.zeros(nl.value, np.int32)
    bl = np.zeros(fo.value, np.int32)
    bo = np.zeros(fo.value, np.int32)
    assert L.nm_mesh_gs_map(nel, lx, p.ctypes.data, fn.ctypes.data, dg.ctypes.data,
                            gd.ctypes.data, bl.ctypes.data, bo.ctypes.data,
                            C.byref(nl), C.byref(fo), C.byref(nb)) == 0
    assert (nl.value, fo.value, nb.value) == (c.nlocal, c.facet_offset, c.nblks)
    assert np.array_equal(dg, c.dg) and np.array_equal(gd, c.gd)
    assert np.array_equal(bl[:nb.value], c.blk_len)
    assert np.array_equal(bo[:nb.value], c.blk_off)
    m = L.nm_mesh_mask_size(nel, lx, fn.ctypes.data)
    msk = np.zeros(max(m, 1), np.int32)
    nm = C.c_int64()
    assert L.nm_mesh_mask(nel, lx, fn.ctypes.data, msk.ctypes.data, C.byref(nm)) == 0
    assert nm.value == c.nmsk
    assert np.array_equal(msk[:nm.value], c.msk[:c.nmsk])


def test_product_rejects_facet_held_three_times():
    from neko_amd._lib import lib
    L = lib()
    # three elements stacked onto the same facet ids
    pid = np.array([[1, 2, 3, 4, 5, 6, 7, 8], [5, 6, 7, 8, 9, 10, 11, 12],
                    [5, 6, 7, 8, 13, 14, 15, 16]], dtype=np.int64)
    fn = np.zeros(18, dtype=np.int32)
    assert L.nm_mesh_facet_neigh(3, pid.ctypes.data, fn.ctypes.data) != 0
    assert b"held by 3" in L.nm_last_error()


@pytest.mark.parametrize("name", ["hemi.nmsh", "lid.nmsh", "poisson_128.nmsh", "tgv_512.nmsh"])
def test_zone_records_match_oracle_reader(name):
    """nm_nmsh_zones: every zone record's element (mapped to file order),
    facet, type and label equal the oracle reader's raw records."""
    from neko_amd.solve import read_nmsh, read_nmsh_zones, zone_facets
    path = os.path.join(GOLD, name)
    z = read_nmsh_zones(path)
    _, _, zo, _ = O.read_nmsh(path)
    _, _, el = read_nmsh(path)
    assert z.shape[0] == zo.shape[0]
    if z.shape[0]:
        assert np.array_equal(el[z[:, 0]], zo[:, 0])
        assert np.array_equal(z[:, 1], zo[:, 1])
        assert np.array_equal(z[:, 2], zo[:, 8])
        assert np.array_equal(z[:, 3], zo[:, 3])
    # labelled facets are the ones the reference reader marks (type 7 only)
    assert len(zone_facets(z)) == int((zo[:, 8] == 7).sum()) if z.shape[0] else True
