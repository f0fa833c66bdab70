"""bigfile block I/O (JERALD/maputils.py:26-76, JERALD/runJERALD.py:47-62).

The bigfile package is absent here, so these tests pin the restated layout
with hand-computed header values, a block typed out by hand in the published
format, and round trips; byte-level agreement with files the library itself
wrote is parity unpinned."""
import os

import numpy as np
import pytest
import torch

from jerald_amd import bigfile_io as B


class _PM:
    def __init__(self, nmesh, box):
        self.Nmesh = np.array(nmesh, dtype=np.int64)
        self.BoxSize = np.array(box, dtype=np.float64)


def test_header_known_answer(tmp_path):
    a = np.arange(6, dtype="<f4").reshape(2, 3)
    B.write_block(str(tmp_path / "blk"), a, nfile=2)
    # byte sums by hand: 1.0f = 00 00 80 3F, 2.0f = 00 00 00 40 -> 0x80+0x3F+0x40 = 255;
    # 3.0f = ..40 40, 4.0f = ..80 40, 5.0f = ..A0 40 -> 0x80+0xC0+0xE0 = 544
    assert (tmp_path / "blk" / "header").read_text() == (
        "DTYPE: <f4\nNMEMB: 3\nNFILE: 2\n000000: 1 : 255 : 255\n000001: 1 : 544 : 544\n")
    assert (tmp_path / "blk" / "000001").read_bytes() == a[1].tobytes()
    # the 16-bit fold of a sum above 2^16: 0x12345 -> 0x2345 + 0x1 = 0x2346
    assert B.checksums(bytes([0xFF] * 293))[1] == ((255 * 293) & 0xFFFF) + ((255 * 293) >> 16)


def test_reads_a_hand_written_block(tmp_path):
    blk = tmp_path / "snap" / "1" / "Position"
    os.makedirs(blk)
    pos = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.5, 8.5, 9.5]], dtype="<f4")
    pos[:2].tofile(blk / "000000")
    pos[2:].tofile(blk / "000001")
    (blk / "header").write_text("DTYPE: <f4\nNMEMB: 3\nNFILE: 2\n000000: 2 : 0 : 0\n000001: 1 : 0 : 0\n")
    # 100.0 as <f8 is 00 00 00 00 00 00 59 40; 256 as <i8 is 00 01 00 ...
    (blk / "attr-v2").write_text("BoxSize <f8 3 000000000000594000000000000059400000000000005940 #HUMANE [ 100.0 100.0 100.0 ]\n"
                                 "Nmesh <i8 1 0001000000000000\n"
                                 "Kind |S1 3 646D6F\n")
    b = B.BigBlock(str(blk))
    assert (b.size, b.nmemb, b.dtype.str) == (3, 3, "<f4")
    np.testing.assert_array_equal(b.attrs["BoxSize"], [100.0] * 3)
    assert b.attrs["Nmesh"].tolist() == [256] and b.attrs["Kind"] == "dmo"
    np.testing.assert_array_equal(b[:], pos)
    np.testing.assert_array_equal(b[1:3], pos[1:3])  # across the file boundary
    got = B.load_positions(str(tmp_path / "snap" / "1"), device="cpu", dtype=torch.float64)
    assert got.dtype == torch.float64
    np.testing.assert_array_equal(got.numpy(), pos.astype(np.float64))


@pytest.mark.parametrize("dt,nfile", [(np.float32, 1), (np.float64, 3)])
def test_savefield_loadfield_round_trip(tmp_path, dt, nfile):
    rng = np.random.default_rng(5)
    m = rng.standard_normal((6, 4, 10)).astype(dt)
    pm = _PM([6, 4, 10], [60.0, 40.0, 100.0])
    B.savefield(m, str(tmp_path / "map"), pm, nfile=nfile)
    out, nmesh, box = B.loadfield(str(tmp_path / "map"))
    assert out.dtype == dt
    np.testing.assert_array_equal(out, m)
    assert nmesh.tolist() == [6, 4, 10] and box.tolist() == [60.0, 40.0, 100.0]
    attrs = B.BigBlock(str(tmp_path / "map" / "Field")).attrs
    assert attrs["ndarray.shape"].tolist() == [6, 4, 10]
    lines = (tmp_path / "map" / "Field" / "attr-v2").read_text().splitlines()
    assert lines[0].startswith("ndarray.shape <i8 3 0600000000000000") and "#HUMANE [ 6 4 10 ]" in lines[0]


def test_rejects_inconsistent_headers(tmp_path):
    blk = tmp_path / "b"
    os.makedirs(blk)
    (blk / "header").write_text("DTYPE: <f4\nNMEMB: 1\nNFILE: 2\n000000: 2 : 0 : 0\n")
    with pytest.raises(ValueError):
        B.BigBlock(str(blk))
    with pytest.raises(ValueError):
        B.write_block(str(tmp_path / "c"), np.zeros((2, 2, 2)))


@pytest.mark.gpu
def test_load_positions_streams_to_gpu(tmp_path):
    rng = np.random.default_rng(3)
    pos = (rng.random((100_003, 3)) * 205.0).astype("<f4")
    B.write_block(str(tmp_path / "1" / "Position"), pos, nfile=4)
    got = B.load_positions(str(tmp_path / "1"), device="cuda", chunk_rows=7_777)
    assert got.device.type == "cuda" and got.dtype == torch.float32
    assert torch.equal(got.cpu(), torch.from_numpy(pos))
    got64 = B.load_positions(str(tmp_path / "1"), device="cuda", dtype=torch.float64, chunk_rows=50_000)
    assert torch.equal(got64.cpu(), torch.from_numpy(pos.astype(np.float64)))


@pytest.mark.gpu
@pytest.mark.parametrize("target", ["Mstar", "HI"])
def test_train_cli_reads_and_writes_bigfile(tmp_path, target):
    """runJERALD-style run on bigfile inputs: FastPM-like Position block,
    target Field (and a star map for HI); writes the best parameters and the
    evaluated map as a Field."""
    import subprocess
    import sys

    from oracle import ldl_oracle as O
    N, L = 16, 32.0
    X = (O.lattice_positions(N, L) + O.zeldovich_displacement(N, L, 1.5, 3)) % L
    B.write_block(str(tmp_path / "snap" / "1" / "Position"), X.astype("<f4"), nfile=2)
    rng = np.random.default_rng(1)
    pm = _PM([N] * 3, [L] * 3)
    B.savefield((rng.random((N, N, N)) + 0.5).astype(np.float32) * (4e-13 if target == "HI" else 1.0),
                str(tmp_path / "ref"), pm)
    args = ["--Nbodypath", str(tmp_path / "snap" / "1"), "--refpath", str(tmp_path / "ref")]
    if target == "HI":
        B.savefield(rng.random((N, N, N)).astype(np.float32), str(tmp_path / "mstar"), pm)
        args += ["--Mstarpath", str(tmp_path / "mstar")]
    r = subprocess.run([sys.executable, "-m", "jerald_amd.train", "--target", target, "--Nmesh", str(N), "--BoxSize", str(L),
                        "--Nopt", "2", "--outpath", str(tmp_path), *args], capture_output=True, text=True, timeout=300,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 0, r.stderr[-2000:]
    stem = tmp_path / f"{target}_Nmesh{N}_Nstep2_n0.00"
    best = np.loadtxt(f"{stem}_bestparam.txt")
    assert best.shape == ((32,) if target == "HI" else (14,))
    m, nmesh, box = B.loadfield(f"{stem}_map")
    assert m.shape == (N, N, N) and np.all(np.isfinite(m)) and np.all(m >= 1e-12)
