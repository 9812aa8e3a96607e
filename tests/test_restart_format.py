"""Restart file framing (CPU only): the library's record writer/reader and the
oracle's restatement against files a Fortran program writes and reads through its
own runtime (amdflang), the subrecord split against the documented layout, and the
reader's errors."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle.restart as orst

FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")

# Writes (mode "write") or reads back and checks (mode "read") four records: a
# scalar, a 3x4 array, a 2x3x2x2 array and a zero-size array, in the way
# RESTART_OUT writes its records (one unformatted WRITE per array).
FORTRAN_KAT = """
program records_kat
  implicit none
  real(8) :: t, a(3,4), b(2,3,2,2), e(0)
  real(8) :: t2, a2(3,4), b2(2,3,2,2), e2(0)
  integer :: i, j, k, l, u
  character(len=512) :: mode, path
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  t = 1.25d0
  do j = 1, 4
    do i = 1, 3
      a(i,j) = i + 10*j + 0.5d0
    end do
  end do
  do l = 1, 2
    do k = 1, 2
      do j = 1, 3
        do i = 1, 2
          b(i,j,k,l) = -(i + 10*j + 100*k + 1000*l) * 1.0d-300
        end do
      end do
    end do
  end do
  if (trim(mode) == 'write') then
    open(newunit=u, file=trim(path), form='UNFORMATTED', status='REPLACE', action='WRITE')
    write(u) t
    write(u) a
    write(u) b
    write(u) e
    close(u)
  else
    open(newunit=u, file=trim(path), form='UNFORMATTED', status='OLD', action='READ')
    read(u) t2
    read(u) a2
    read(u) b2
    read(u) e2
    close(u)
    if (t2 == t .and. all(a2 == a) .and. all(b2 == b)) then
      print '(A)', 'READ OK'
    else
      print '(A)', 'READ MISMATCH'
    end if
  end if
end program records_kat
"""


def kat_arrays():
    a = np.empty((3, 4), order="F")
    for j in range(4):
        for i in range(3):
            a[i, j] = (i + 1) + 10 * (j + 1) + 0.5
    b = np.empty((2, 3, 2, 2), order="F")
    for idx in np.ndindex(b.shape):
        i, j, k, l = (x + 1 for x in idx)
        b[idx] = -(i + 10 * j + 100 * k + 1000 * l) * 1.0e-300
    return [np.array([1.25]), a, b, np.zeros(0)]


@pytest.fixture(scope="module")
def fortran_prog(tmp_path_factory):
    if not (os.path.exists(FC) or shutil.which(FC)):
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("fkat")
    src = d / "kat.f90"
    src.write_text(FORTRAN_KAT)
    exe = d / "kat"
    subprocess.check_call([FC, "-O0", str(src), "-o", str(exe)], cwd=d)
    return str(exe)


def test_fortran_file_known_answer(trt, fortran_prog, tmp_path):
    """Bytes the Fortran runtime writes == the library's writer == the oracle's framing;
    the library reads the Fortran file and the Fortran program reads the library's."""
    ffile, lfile = tmp_path / "fortran.bin", tmp_path / "lib.bin"
    subprocess.check_call([fortran_prog, "write", str(ffile)])
    arrs = kat_arrays()
    trt.native.write_records(str(lfile), arrs)
    fbytes = ffile.read_bytes()
    assert lfile.read_bytes() == fbytes
    assert b"".join(orst.frame(a.tobytes(order="F")) for a in arrs) == fbytes
    # single-subrecord framing: 4-byte little-endian lengths around each record
    assert fbytes[:4] == struct.pack("<i", 8) and len(fbytes) == sum(8 + 8 * a.size for a in arrs)
    got = trt.native.read_records(str(ffile), [a.size for a in arrs])
    for g, a in zip(got, arrs):
        assert np.array_equal(g, a.ravel(order="F"))
    out = subprocess.check_output([fortran_prog, "read", str(lfile)], text=True)
    assert "READ OK" in out


def test_split_records_follow_documented_layout(trt, tmp_path, monkeypatch):
    """A 40-byte record at a 16-byte subrecord limit: leading markers negated while more
    follow, trailing markers negated after the first subrecord."""
    d = np.arange(1.0, 6.0)
    raw = d.tobytes()
    p = struct.pack
    want = (p("<i", -16) + raw[0:16] + p("<i", 16) + p("<i", -16) + raw[16:32] + p("<i", -16) +
            p("<i", 8) + raw[32:40] + p("<i", -8))
    assert orst.frame(raw, 16) == want
    monkeypatch.setenv("TRT_SUBRECORD_LENGTH", "16")
    f = tmp_path / "split.bin"
    trt.native.write_records(str(f), [d])
    assert f.read_bytes() == want
    # an exact multiple: no empty trailing subrecord
    trt.native.write_records(str(f), [d[:4]])
    assert f.read_bytes() == p("<i", -16) + raw[0:16] + p("<i", 16) + p("<i", 16) + raw[16:32] + p("<i", -16)


@pytest.mark.parametrize("limit", ["12", "8", "1", "2147483639"])
def test_split_round_trip(trt, tmp_path, monkeypatch, limit):
    """Splits that cut doubles in two read back exactly, by the library and the oracle."""
    rng = np.random.default_rng(int(limit) % 97)
    arrs = [rng.standard_normal(n) for n in (1, 7, 0, 33)]
    monkeypatch.setenv("TRT_SUBRECORD_LENGTH", limit)
    f = tmp_path / "rt.bin"
    trt.native.write_records(str(f), arrs)
    assert f.read_bytes() == b"".join(orst.frame(a.tobytes(), int(limit)) for a in arrs)
    monkeypatch.delenv("TRT_SUBRECORD_LENGTH")
    got = trt.native.read_records(str(f), [a.size for a in arrs])
    buf, pos = f.read_bytes(), 0
    for g, a in zip(got, arrs):
        assert np.array_equal(g, a)
        data, pos = orst.unframe(buf, pos)
        assert data == a.tobytes()
    assert pos == len(buf)


def test_append_and_time_record(trt, tmp_path):
    f = tmp_path / "t.bin"
    trt.native.write_records(str(f), [np.array([3.5])])
    trt.native.write_records(str(f), [np.arange(4.0)], append=True)
    assert trt.native.restart_time_in(str(f)) == 3.5
    t, a = trt.native.read_records(str(f), [1, 4])
    assert t[0] == 3.5 and np.array_equal(a, np.arange(4.0))


def test_reader_errors(trt, tmp_path, monkeypatch):
    nat = trt.native
    f = tmp_path / "e.bin"
    nat.write_records(str(f), [np.arange(6.0), np.arange(3.0)])
    with pytest.raises(nat.TrtError, match=r"record 2 .*holds 24 bytes, expected 32"):
        nat.read_records(str(f), [6, 4])
    with pytest.raises(nat.TrtError, match=r"record 1 .*holds 48 bytes, expected 40"):
        nat.read_records(str(f), [5])
    with pytest.raises(nat.TrtError, match="ends before the record"):
        nat.read_records(str(f), [6, 3, 1])
    raw = f.read_bytes()
    (tmp_path / "cut.bin").write_bytes(raw[:-10])
    with pytest.raises(nat.TrtError, match="ends inside the record"):
        nat.read_records(str(tmp_path / "cut.bin"), [6, 3])
    bad = bytearray(raw)
    bad[4 + 48:4 + 48 + 4] = struct.pack("<i", 40)        # trailing marker of record 1
    (tmp_path / "bad.bin").write_bytes(bytes(bad))
    with pytest.raises(nat.TrtError, match="markers disagree"):
        nat.read_records(str(tmp_path / "bad.bin"), [6])
    with pytest.raises(nat.TrtError, match="could not be opened"):
        nat.read_records(str(tmp_path / "missing.bin"), [1])
    with pytest.raises(nat.TrtError, match="could not be opened"):
        nat.restart_time_in(str(tmp_path / "missing.bin"))
    monkeypatch.setenv("TRT_SUBRECORD_LENGTH", "0")
    with pytest.raises(nat.TrtError, match="TRT_SUBRECORD_LENGTH"):
        nat.write_records(str(f), [np.arange(2.0)])
    assert f.read_bytes() == raw and not (tmp_path / "e.bin.tmp").exists()


def test_oracle_restart_arrays_order_and_rt_src(oracle_mod, tmp_path):
    """restart_arrays follows RESTART_OUT's record list and expands RT_Src over the
    corners and directions; write_restart/read_restart round-trip it."""
    import importlib
    trt = importlib.import_module("2d_trt_rom_system_amd")
    prob = trt.problem.small(nx=5, ny=4, order=4, ng=2)
    o = oracle_mod.Oracle(prob, threads=1)
    o.step()
    arrs = orst.restart_arrays(o, prob.nm)
    assert len(arrs) == 25 and orst.RECORDS[2] == "I_crn" and orst.RECORDS[9] == "RT_Src"
    rt = arrs[9]
    assert rt.shape == (2 * prob.nx, 2 * prob.ny, prob.nm, prob.ng)
    src = np.asarray(o["RT_Src"])
    for (ci, cj, m, g) in [(0, 0, 0, 0), (1, 1, prob.nm - 1, 1), (2 * prob.nx - 1, 2 * prob.ny - 1, 3, 0)]:
        assert rt[ci, cj, m, g] == src[ci // 2, cj // 2, g]
    assert arrs[0][0] == o.time and arrs[2].shape == (2 * prob.nx, 2 * prob.ny, prob.nm, prob.ng)
    f = str(tmp_path / "o.rst")
    orst.write_restart(f, arrs)
    back = orst.read_restart(f, [a.shape for a in arrs])
    for a, b in zip(arrs, back):
        assert np.array_equal(a, b)


def test_failed_write_keeps_the_existing_file(trt, tmp_path):
    """A write that fails part-way (here: record 2 has no data) leaves the file that was
    there untouched and no temporary file behind."""
    import ctypes as C
    nat = trt.native
    f = tmp_path / "keep.bin"
    nat.write_records(str(f), [np.arange(3.0)])
    before = f.read_bytes()
    a = np.arange(5.0)
    dp = C.POINTER(C.c_double)
    ptrs = (dp * 2)(a.ctypes.data_as(dp), dp())
    cnt = (C.c_size_t * 2)(5, 4)
    assert nat.load().trt_frecords_write(str(f).encode(), 2, ptrs, cnt, 0) != 0
    assert "record 2 has no data" in nat.load().trt_frecords_last_error().decode()
    assert f.read_bytes() == before
    assert sorted(p.name for p in tmp_path.iterdir()) == ["keep.bin"]
