"""Curved elements (Neko's .nmsh curve records; dofmap_generate_xyz,
src/sem/dofmap.f90:797-1285) on the CPU.

* The reader: nm_nmsh_curves against the oracle's restatement of the
  packed 532-byte record (nmsh_file.f90:575-639, format/nmsh.f90:39-43) on
  the reference's cylinder mesh.
* The library's per-element coordinates (nm_curve_xyz, a closed form of
  the Gordon-Hall midpoint map plus the arc perturbation) against the
  oracle's step-by-step restatement (or_curve_element), on every curved
  element of that mesh and on synthetic elements with arcs and midpoints
  on every edge, both signs of the radius.
* The oracle pinned by what the geometry fixes:
  - a midpoint on the straight edge leaves the trilinear map unchanged;
  - the cylinder mesh's volume equals the closed form of a box minus a
    prism over a 16-gon with parabolic caps on its sides (the quadratic
    midpoint edges), and the straight mesh's the box minus the 16-gon
    prism;
  - an element with one midpoint edge, extruded in z, carries at its centre
    half the edge's midpoint deviation (2-D transfinite interpolation);
  - an annular-sector element with concentric arcs on its inner and outer
    r-edges maps every GLL point onto the circle of radius
    (1-s)/2 R_in + (1+s)/2 R_out at the polar angle of its r coordinate.
No GPU is touched."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CYL = os.path.join(ROOT, "tests", "golden", "nmsh", "small_test_cyl.nmsh")


def _lib_or_skip():
    from neko_amd import _lib
    try:
        return _lib.lib()
    except OSError:
        pytest.skip("libneko_amd.so not built")


def test_fixture_is_the_reference_file():
    ref = "/root/reference/tests/integration/meshes/small_test_cyl.nmsh"
    if not os.path.exists(ref):
        pytest.skip("reference checkout not present")
    a = hashlib.sha256(open(CYL, "rb").read()).hexdigest()
    assert a == hashlib.sha256(open(ref, "rb").read()).hexdigest()


def test_reader_matches_oracle():
    _lib_or_skip()
    from neko_amd import solve as S
    el, data, ty = S.read_nmsh_curves(CYL)
    el_o, data_o, ty_o = O.read_nmsh_curves(CYL)
    assert len(el) == 48
    assert np.array_equal(el, el_o) and np.array_equal(ty, ty_o)
    assert np.array_equal(data, data_o)
    # the file's curves: midpoints (type 4) on the bottom and top r-edges,
    # all on the cylinder of radius 0.5
    assert set(np.unique(ty)) == {0, 4}
    mids = data.reshape(-1, 12, 5)[ty == 4]
    assert np.allclose(np.hypot(mids[:, 0], mids[:, 1]), 0.5, atol=1e-8)
    # vertices are still read (straight), the curves come separately
    v, pid, _ = S.read_nmsh(CYL)
    assert v.shape == (336, 8, 3)


def _synthetic(seed):
    """A perturbed element around the origin's quadrant with random curve
    types: arcs (3) of random sign on edges 1-8, midpoints (4) anywhere."""
    rng = np.random.default_rng(seed)
    v = np.array([[cx, cy, cz] for cz in (0, 1) for cy in (0, 1) for cx in (0, 1)], float)
    v = v + 0.05 * rng.standard_normal(v.shape)
    ty = np.zeros(12, np.int32)
    cd = np.zeros((12, 5))
    for e in range(12):
        u = rng.random()
        if e < 8 and u < 0.35:
            ty[e] = 3
            cd[e, 0] = rng.choice([-1, 1]) * rng.uniform(1.5, 4.0)
        elif u < 0.7:
            ty[e] = 4
            cd[e, :3] = rng.uniform(-0.2, 1.2, 3)
    return v, cd, ty


@pytest.mark.parametrize("lx", [2, 3, 5, 8])
def test_library_element_coordinates_match_oracle(lx):
    L = O.lib()
    _lib_or_skip()
    from neko_amd import solve as S
    z, _ = O.zwgll(lx)
    n3 = lx ** 3
    cases = []
    v, pid, _, _ = O.read_nmsh(CYL)
    el, data, ty = O.read_nmsh_curves(CYL)
    for c in range(len(el)):
        cases.append((v[el[c]], data[c], ty[c]))
    for seed in range(40):
        cases.append(_synthetic(seed))
    worst = 0.0
    for ve, cd, t in cases:
        ve = np.ascontiguousarray(ve, float).reshape(24)
        cd = np.ascontiguousarray(cd, float).reshape(60)
        t = np.ascontiguousarray(t, np.int32).reshape(12)
        ref = [np.zeros(n3) for _ in range(3)]
        L.or_xyz(1, lx, z, ve, *ref)
        assert L.or_curve_element(lx, z, ve, cd, t, *ref) == 0
        got = S.curve_xyz(lx, ve, cd, t)
        scale = max(1.0, np.max(np.abs(ref)))
        worst = max(worst, max(np.max(np.abs(got[c] - ref[c])) for c in range(3)) / scale)
    assert worst < 1e-13, worst


# ------------------------------------------------------------ oracle pins -
def _element_xyz(lx, v, cd, t):
    L = O.lib()
    z, _ = O.zwgll(lx)
    out = [np.zeros(lx ** 3) for _ in range(3)]
    ve = np.ascontiguousarray(v, float).reshape(24)
    L.or_xyz(1, lx, z, ve, *out)
    assert L.or_curve_element(lx, z, ve, np.ascontiguousarray(cd, float).reshape(60),
                              np.ascontiguousarray(t, np.int32).reshape(12), *out) == 0
    return np.array(out), z


CYC = [(0, 1), (1, 3), (3, 2), (2, 0), (4, 5), (5, 7), (7, 6), (6, 4),
       (0, 4), (1, 5), (3, 7), (2, 6)]


def test_oracle_straight_midpoints_leave_trilinear_map():
    rng = np.random.default_rng(4)
    v = np.array([[cx, cy, cz] for cz in (0, 1) for cy in (0, 1) for cx in (0, 1)], float)
    v += 0.1 * rng.standard_normal(v.shape)
    cd = np.zeros((12, 5))
    for e, (a, b) in enumerate(CYC):
        cd[e, :3] = 0.5 * (v[a] + v[b])
    t = np.full(12, 4, np.int32)
    for lx in (3, 6):
        got, _ = _element_xyz(lx, v, cd, t)
        lin, _ = _element_xyz(lx, v, cd, np.zeros(12, np.int32))
        assert np.max(np.abs(got - lin)) < 1e-14


def test_oracle_cylinder_volume_closed_form():
    v, pid, _, _ = O.read_nmsh(CYL)
    cur = O.read_nmsh_curves(CYL)
    R, H, nseg = 0.5, 6.0, 16
    box = 40.0 * 30.0 * H
    poly = 0.5 * nseg * R * R * np.sin(2 * np.pi / nseg)
    chord = 2 * R * np.sin(np.pi / nseg)
    sag = R * (1 - np.cos(np.pi / nseg))
    caps = nseg * (2.0 / 3.0) * chord * sag  # parabolic segments
    for lx in (3, 5):
        curved = O.OracleCase.from_mesh(v, pid, lx, curves=cur)
        straight = O.OracleCase.from_mesh(v, pid, lx)
        assert abs(curved.B.sum() - (box - H * (poly + caps))) < 1e-9 * box
        assert abs(straight.B.sum() - (box - H * poly)) < 1e-9 * box
        # and the curved mesh is the closer one to the true cylinder
        exact = box - H * np.pi * R * R
        assert abs(curved.B.sum() - exact) < 0.01 * abs(straight.B.sum() - exact)


def test_oracle_midpoint_edge_transfinite_centre():
    v = np.array([[cx * 2.0, cy * 1.0, cz * 3.0] for cz in (0, 1) for cy in (0, 1)
                  for cx in (0, 1)], float)
    cd = np.zeros((12, 5))
    t = np.zeros(12, np.int32)
    dev = np.array([0.1, -0.3, 0.2])  # deviation of edges 1 (t=-1) and 5 (t=+1)
    for e in (0, 4):
        a, b = CYC[e]
        cd[e, :3] = 0.5 * (v[a] + v[b]) + dev
        t[e] = 4
    lx = 5  # odd: a GLL point at the centre
    xyz, z = _element_xyz(lx, v, cd, t)
    c = (lx // 2) * (1 + lx + lx * lx)
    centre = v.mean(axis=0)
    assert np.max(np.abs(xyz[:, c] - (centre + 0.5 * dev))) < 1e-14
    # the face s = +1 (j = lx-1) is untouched
    lin, _ = _element_xyz(lx, v, cd, np.zeros(12, np.int32))
    j1 = np.array([i + lx * (lx - 1) + lx * lx * k for k in range(lx) for i in range(lx)])
    assert np.max(np.abs(xyz[:, j1] - lin[:, j1])) < 1e-14


@pytest.mark.parametrize("sign", [1, -1])
def test_oracle_annular_sector_is_polar(sign):
    """Inner edge 1 (v1 -> v2) and outer edge 3 (v4 -> v3), bottom and top,
    on concentric circles about the origin; the sign of each radius is the
    one that puts the arc's centre at the origin, and the mirrored element
    (x -> -x) takes the opposite signs."""
    R0, R1, th_a, th_b = 1.0, 2.0, 0.2, 0.9
    pts = {}
    for cy, R in ((0, R0), (1, R1)):
        for cx, th in ((0, th_b), (1, th_a)):
            pts[(cx, cy)] = (R * np.cos(th), R * np.sin(th))
    v = np.array([[*pts[(cx, cy)], float(cz)] for cz in (0, 1) for cy in (0, 1)
                  for cx in (0, 1)])
    cd = np.zeros((12, 5))
    t = np.zeros(12, np.int32)
    # the sign of the radius puts the centre on one side of the chord: pick
    # for each edge the sign that puts it at the origin
    for e, R in ((0, R0), (2, R1), (4, R0), (6, R1)):
        a, b = CYC[e]
        p1, p2 = v[a, :2], v[b, :2]
        n = np.array([p2[1] - p1[1], p1[0] - p2[0]])
        mid = 0.5 * (p1 + p2)
        s = 1.0 if np.dot(-n, -mid) > 0 else -1.0
        cd[e, 0] = s * R
        t[e] = 3
    if sign < 0:
        # the mirrored element (x -> -x) needs the opposite signs
        v[:, 0] *= -1
        cd[:, 0] *= -1
    lx = 7
    xyz, z = _element_xyz(lx, v, cd, t)
    r = np.hypot(xyz[0], xyz[1])
    for k in range(lx):
        for j in range(lx):
            for i in range(lx):
                p = i + lx * j + lx * lx * k
                want = (1 - z[j]) / 2 * R0 + (1 + z[j]) / 2 * R1
                assert abs(r[p] - want) < 1e-13, (i, j, k, r[p], want)
    # and the polar angle depends on the r coordinate alone
    ang = np.arctan2(xyz[1], xyz[0]).reshape(lx, lx, lx)  # [k, j, i]
    assert np.max(np.abs(ang - ang[:1, :1, :])) < 1e-13
