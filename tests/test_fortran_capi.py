"""The Fortran interface module neko_amd/fortran/neko_amd_capi.f90 (the
iso_c_binding layer Neko's plug-ins use) compiles with the image's flang,
links against libneko_amd.so, and a Fortran program calling host entry
points through it gets the same results as the C ABI called from Python:
the GLL rule and D (nm_gll) and Neko's gs maps of a periodic box
(nm_box_gs_sizes / nm_box_gs_map).  No GPU is touched."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/llvm/bin/flang"
LIBDIR = os.path.join(ROOT, "neko_amd")

PROGRAM = r"""
program capi_check
  use, intrinsic :: iso_c_binding
  use neko_amd_capi
  implicit none
  integer, parameter :: lx = 6
  real(c_double) :: z(lx), w(lx), d(lx*lx), dt(lx*lx)
  integer(c_int64_t) :: nlocal, facet_offset, nblks
  integer(c_int), allocatable :: dg(:), gd(:), bl(:), bo(:)
  integer :: st, i
  st = nm_gll(lx, z, w, d, dt)
  if (st /= 0) stop 1
  write(*, '(a)') 'GLL'
  do i = 1, lx
     write(*, '(3es26.17)') z(i), w(i), d(i + lx)
  end do
  st = nm_box_gs_sizes(3, 2, 3, 4, 1, 0, 0, nlocal, facet_offset)
  if (st /= 0) stop 2
  allocate(dg(nlocal), gd(nlocal), bl(facet_offset), bo(facet_offset))
  st = nm_box_gs_map(3, 2, 3, 4, 1, 0, 0, dg, gd, bl, bo, nlocal, &
       facet_offset, nblks)
  if (st /= 0) stop 3
  write(*, '(a, 3i12)') 'GS', nlocal, facet_offset, nblks
  write(*, '(i12)') sum(int(dg, c_int64_t) * int(gd, c_int64_t))
  write(*, '(i12)') sum(int(bl(1:nblks), c_int64_t) * (/ (int(i, c_int64_t), i = 1, int(nblks)) /))
end program capi_check
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="no flang in this image")
def test_fortran_capi_compiles_links_and_agrees(tmp_path):
    from neko_amd import solve as S
    from neko_amd._lib import LIB_PATH
    assert os.path.exists(LIB_PATH)
    src = tmp_path / "prog.f90"
    src.write_text(PROGRAM)
    shutil.copy(os.path.join(ROOT, "neko_amd", "fortran", "neko_amd_capi.f90"), tmp_path)
    r = subprocess.run([FC, "-c", "neko_amd_capi.f90"], cwd=tmp_path, capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([FC, "prog.f90", "neko_amd_capi.o", "-o", "prog", f"-L{LIBDIR}",
                        "-lneko_amd", f"-Wl,-rpath,{LIBDIR}"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(tmp_path / "prog")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    i = lines.index("GLL")
    vals = np.array([[float(x) for x in lines[i + 1 + k].split()] for k in range(6)])
    sp = S.Space(6)
    assert np.array_equal(vals[:, 0], sp.zg) and np.array_equal(vals[:, 1], sp.wx)
    assert np.array_equal(vals[:, 2], sp.dx[6:12])
    j = [k for k, ln in enumerate(lines) if ln.startswith("GS")][0]
    nl, fo, nb = (int(x) for x in lines[j].split()[1:])
    m = S.box_gs_map(3, 2, 3, 4, periodic=(1, 0, 0))
    assert (nl, fo, nb) == (m["nlocal"], m["facet_offset"], m["nblks"])
    assert int(lines[j + 1]) == int(np.sum(m["dg"].astype(np.int64) * m["gd"]))
    assert int(lines[j + 2]) == int(np.sum(m["blk_len"].astype(np.int64) *
                                           np.arange(1, nb + 1)))
