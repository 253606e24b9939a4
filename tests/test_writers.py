"""engstore's and engconst's text writers against the Fortran runtime.

A small Fortran program, compiled here with the same compiler as the host
driver, reads values from a stream file and writes them with the format
statements of erdst/engproc.F90 (engfile :683-693, weights :641-656, aveuv
:631-639, uvrange :658-669, slfeng :712-722, flcuv :485-499); the Python
writers must reproduce its output byte for byte on the same values, including
the real (float32) storage of aveuv, minuv/maxuv, flceng and the self-energy
bin coordinate.  The values cover both branches of G editing, E editing,
negative zero after rounding, exact decimal ties and the 1e5 switch of
uvrange.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from erdst_gpu import engstore

FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")

PROGRAM = r"""
program fmt
  implicit none
  integer :: n, nsl, i, u
  real(8), allocatable :: d(:), e(:)
  real, allocatable :: r(:,:)
  open(newunit=u, file='vals.bin', access='stream', form='unformatted', status='old')
  read(u) n, nsl
  allocate(d(n), e(n + 1), r(nsl, n))
  read(u) d, e, r
  close(u)
  open(newunit=u, file='eng.txt', action='write')
  do i = 1, n
     write(u, '(g15.7,g15.7,i5,g25.15,g25.15)') d(i), e(i), i, d(i) * 3.0d0, e(i) / 7.0d0
  end do
  close(u)
  open(newunit=u, file='weight.txt', action='write')
  do i = 1, n
     write(u, '(i5,e20.8)') i, d(i)
     write(u, '(i5,e20.8,e20.7)') i, d(i), e(i)
  end do
  close(u)
  open(newunit=u, file='aveuv.txt', action='write')
  do i = 1, n
     write(u, '(i5,9999f15.5)') i, real(d(i)), real(e(i))
  end do
  close(u)
  open(newunit=u, file='uvrange.txt', action='write')
  write(u, '(A)') ' species     minimum        maximum'
  do i = 1, n
     if (real(e(i)) < 1.0e5) then
        write(u, '(i5,2f15.5)') i - 1, real(d(i)), real(e(i))
     else
        write(u, '(i5,f15.5,g18.5)') i - 1, real(d(i)), real(e(i))
     end if
  end do
  close(u)
  open(newunit=u, file='slfeng.txt', action='write')
  do i = 1, n
     if (i < n) then
        write(u, '(g15.7,g25.15)') real((e(i) + e(i + 1)) / 2.0), d(i)
     else
        write(u, '(g15.7,g25.15)') real(e(n)), d(i)
     end if
  end do
  close(u)
  open(newunit=u, file='flcuv1.txt', action='write')
  write(u, '(i9, 9999f15.5)') 17, r(1:nsl, 1)
  close(u)
  open(newunit=u, file='flcuvn.txt', action='write')
  do i = 1, n
     write(u, '(2i9, 9999f15.5)') i, 17, r(1:nsl, i)
  end do
  close(u)
  open(newunit=u, file='progress.txt', action='write')
  write(u, *) 17
  write(u, *) 123456
  close(u)
end program fmt
"""


@pytest.fixture(scope="module")
def fortran(tmp_path_factory):
    if shutil.which(FLANG) is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("fmt")
    (d / "fmt.f90").write_text(PROGRAM)
    subprocess.run([FLANG, "-O0", "fmt.f90", "-o", "fmt"], cwd=d, check=True, capture_output=True)
    return d


def _values():
    rng = np.random.default_rng(7)
    d = np.concatenate([
        [0.0, -0.0, 1.0, -1.0, 0.1, 0.09999999, 0.015625, -0.015625, 2.0 ** -7, 9999999.4, 12345678.0,
         -3.2e-6, -4.9e-6, 1.0e-300, 6.02e23, -123.456789, 0.5, 1e5, 99999.99],
        rng.normal(0, 1, 30) * 10.0 ** rng.integers(-8, 9, 30)])
    e = np.concatenate([np.sort(rng.uniform(-20, 20, d.size - 4)), [120.0, 99999.0, 1.0e5, 3.3e7], [4.0e8]])
    return d, e


def _run(fortran, nsl=3):
    d, e = _values()
    n = d.size
    r = (np.random.default_rng(9).normal(0, 1, (n, nsl)) * 10.0 ** np.random.default_rng(8).integers(-3, 6, (n, nsl)))
    r = r.astype(np.float32)
    with open(fortran / "vals.bin", "wb") as fh:
        np.array([n, nsl], np.int32).tofile(fh)
        d.astype(np.float64).tofile(fh)
        e.astype(np.float64).tofile(fh)
        np.asfortranarray(r.T).ravel(order="F").tofile(fh)
    for f in ("eng.txt", "weight.txt", "aveuv.txt", "uvrange.txt", "slfeng.txt", "flcuv1.txt", "flcuvn.txt",
              "progress.txt"):
        p = fortran / f
        if p.exists():
            p.unlink()
    subprocess.run(["./fmt"], cwd=fortran, check=True, capture_output=True)
    return d, e, r


def _read(p):
    return open(p).read()


def test_engfile_weights_aveuv_formats(fortran, tmp_path):
    d, e, _ = _run(fortran)
    n = d.size
    # engfile lines through the writer's own formatter calls
    lines = "".join(engstore.fortran_g(d[i], 15, 7) + engstore.fortran_g(e[i], 15, 7) + f"{i + 1:5d}"
                    + engstore.fortran_g(d[i] * 3.0, 25, 15) + engstore.fortran_g(e[i] / 7.0, 25, 15) + "\n"
                    for i in range(n))
    assert lines == _read(fortran / "eng.txt")
    # weights: the writer emits the two forms in separate files; interleave them
    engstore.write_weights(tmp_path / "w1", d)
    engstore.write_weights(tmp_path / "w2", d, e[:n])
    w1 = _read(tmp_path / "w1").splitlines()
    w2 = _read(tmp_path / "w2").splitlines()
    got = "".join(a + "\n" + b + "\n" for a, b in zip(w1, w2))
    assert got == _read(fortran / "weight.txt")
    engstore.write_aveuv(tmp_path / "aveuv", np.stack([d, e[:n]], 1))
    assert _read(tmp_path / "aveuv") == _read(fortran / "aveuv.txt")


def test_uvrange_slfeng_formats(fortran, tmp_path):
    d, e, _ = _run(fortran)
    n = d.size
    engstore.write_uvrange(tmp_path / "uvrange", d, e[:n])
    assert _read(tmp_path / "uvrange") == _read(fortran / "uvrange.txt")
    engstore.write_slfeng(tmp_path / "slfeng", d, e[:n])
    assert _read(tmp_path / "slfeng") == _read(fortran / "slfeng.txt")


def test_flcuv_and_progress_formats(fortran):
    d, e, r = _run(fortran)
    one = "".join(ln + "\n" for ln in engstore.flcuv_lines(17, r[:1].astype(np.float64)))
    assert one == _read(fortran / "flcuv1.txt")
    many = "".join(ln + "\n" for ln in engstore.flcuv_lines(17, r.astype(np.float64)))
    assert many == _read(fortran / "flcuvn.txt")
    assert engstore.progress_line(17) + "\n" + engstore.progress_line(123456) + "\n" == _read(fortran / "progress.txt")


def test_block_writer_frame_files(tmp_path):
    uvmax, uvsoft, uvcrd = np.array([3]), np.array([2]), np.array([-1.0, 0.0, 1.0])
    wr = engstore.BlockWriter(tmp_path, uvmax, uvsoft, uvcrd, 0.6, soln=True, engdiv=1)
    wr.frame(np.array([[-1.5, 2.25]]))
    wr.frame(np.array([[0.5, -0.125]]))
    assert _read(tmp_path / "flcuv.tt").splitlines() == [
        f"{1:9d}{-1.5:15.5f}{2.25:15.5f}", f"{2:9d}{0.5:15.5f}{-0.125:15.5f}"]
    wr2 = engstore.BlockWriter(tmp_path, uvmax, uvsoft, uvcrd, 0.6, soln=True, engdiv=1)
    with pytest.raises(FileExistsError):       # status='new'
        wr2.frame(np.array([[0.0, 0.0]]))
    wr3 = engstore.BlockWriter(tmp_path, uvmax, uvsoft, uvcrd, 0.6, soln=False, engdiv=1)
    for _ in range(3):
        wr3.frame(skpcnf=5)
    assert _read(tmp_path / "progress.tt") == " 5\n 10\n 15\n"


def test_selfenergy_bin_coords():
    escrd = np.array([-1.0, -0.5, 0.1, 2.0])
    c = engstore.selfenergy_bin_coords(escrd)
    assert c.dtype == np.float32
    np.testing.assert_array_equal(c, np.float32([-0.75, -0.2, 1.05, 2.0]))


def test_bin_info_without_core_region():
    """A species without a core region: its last soft bin averages with the
    next species' first point (as the reference reads uvcrd(iduv + 1)); the
    last bin of the last species has no next point and is the open top bin."""
    uvmax, uvsoft = np.array([3, 3]), np.array([3, 3])
    uvcrd = np.array([-1.0, 0.0, 1.0, -2.0, -1.5, -1.0])
    left, mid, inv, spec = engstore.bin_info(uvmax, uvsoft, uvcrd)
    np.testing.assert_array_equal(spec, [1, 1, 1, 2, 2, 2])
    assert mid[2] == (1.0 + -2.0) / 2 and inv[2] == 1.0 / (-2.0 - 1.0)
    assert mid[5] == -1.0 and inv[5] == 0.0
    np.testing.assert_array_equal(mid[:2], [-0.5, 0.5])
