"""bigfile blocks: the reference's on-disk format for maps and particles.

The reference reads FastPM particle positions (``File(path)['Position']``,
JERALD/runJERALD.py:47-62) and reads and writes density/target maps as a
``Field`` block (JERALD/maputils.py:26-76) with the ``bigfile`` package
(github.com/rainwoodman/bigfile; imported at maputils.py:3 and
runJERALD.py:10 and pinned nowhere -- setup.py:114-119 does not list it).
The package is not installed here, so this module restates its on-disk
layout (the BigFile directory format of that library) in NumPy:

  <file>/                 a directory; each block is a sub-directory
    <block>/header        text: ``DTYPE: <f4`` / ``NMEMB: 3`` / ``NFILE: n``,
                          then one line per data file,
                          ``%06X: <rows> : <byte sum> : <folded sum>``
    <block>/attr-v2       text, one attribute per line:
                          ``name dtype nmemb HEXBYTES #HUMANE [ values ]``
    <block>/000000 ...    raw rows of NMEMB items of DTYPE, split over the
                          files in order (file names are the index in hex)

``savefield`` / ``loadfield`` mirror maputils.py:26-76 on one process;
``load_positions`` streams a particle block into a GPU tensor through a
pinned staging buffer, so a 1024^3 snapshot (12.9 GB) never sits whole in
host memory.
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["BigBlock", "write_block", "savefield", "loadfield", "load_positions", "checksums"]


def checksums(buf):
    """(byte sum mod 2^32, its 16-bit fold) of a data file, as the header
    records them."""
    b = np.frombuffer(memoryview(buf).cast("B"), dtype=np.uint8)
    s = int(b.sum(dtype=np.uint64)) & 0xFFFFFFFF
    r = (s & 0xFFFF) + (s >> 16)
    return s, (r & 0xFFFF) + (r >> 16)


def _parse_attrs(path):
    attrs = {}
    fn = os.path.join(path, "attr-v2")
    if not os.path.exists(fn):
        return attrs
    with open(fn) as f:
        for line in f:
            tok = line.split()
            if len(tok) < 3:
                continue
            name, dt, nmemb = tok[0], np.dtype(tok[1]), int(tok[2])
            raw = bytes.fromhex(tok[3]) if nmemb > 0 and len(tok) > 3 else b""
            v = np.frombuffer(raw, dtype=dt, count=nmemb)
            attrs[name] = v.tobytes().decode() if dt.kind == "S" else v.copy()
    return attrs


def _humane(v):
    v = np.asarray(v)
    if v.dtype.kind == "S":
        return repr(v.tobytes().decode())
    return " ".join(repr(x.item()) for x in v.reshape(-1))


def _write_attrs(path, attrs):
    lines = []
    for name, val in attrs.items():
        if " " in name:
            raise ValueError("attribute names cannot contain spaces")
        if isinstance(val, (str, bytes)):
            raw = val.encode() if isinstance(val, str) else val
            v = np.frombuffer(raw, dtype="S1")
        else:
            v = np.atleast_1d(np.asarray(val))
            if v.dtype.kind not in "biuf":
                raise ValueError(f"attribute {name}: unsupported type {v.dtype}")
            v = v.reshape(-1)
        lines.append(f"{name} {v.dtype.str} {v.size} {v.tobytes().hex().upper()} #HUMANE [ {_humane(v)} ]\n")
    with open(os.path.join(path, "attr-v2"), "w") as f:
        f.writelines(lines)


class BigBlock:
    """Read side of one block: ``dtype`` (item type), ``nmemb``, ``size``
    (rows), ``attrs`` and row slices ``block[a:b]`` (shape (rows, nmemb), or
    (rows,) when nmemb == 1)."""

    def __init__(self, path):
        self.path = path
        hdr = {}
        files = []
        with open(os.path.join(path, "header")) as f:
            for line in f:
                if not line.strip():
                    continue
                key, _, rest = line.partition(":")
                key = key.strip()
                if key in ("DTYPE", "NMEMB", "NFILE"):
                    hdr[key] = rest.strip()
                else:
                    parts = [p.strip() for p in rest.split(":")]
                    files.append((int(key, 16), int(parts[0])))
        self.dtype = np.dtype(hdr["DTYPE"])
        self.nmemb = int(hdr["NMEMB"])
        nfile = int(hdr["NFILE"])
        files.sort()
        if [i for i, _ in files] != list(range(nfile)):
            raise ValueError(f"{path}/header lists files {[i for i, _ in files]}, expected {nfile}")
        self.fsize = np.array([n for _, n in files], dtype=np.int64)
        self.foff = np.concatenate([[0], np.cumsum(self.fsize)])
        self.size = int(self.foff[-1])
        self.attrs = _parse_attrs(path)

    def _shape(self, n):
        return (n,) if self.nmemb == 1 else (n, self.nmemb)

    def read(self, start=0, stop=None, out=None):
        """Rows [start, stop) into ``out`` (a NumPy array of that many rows,
        any layout-compatible buffer) or a new array."""
        stop = self.size if stop is None else min(stop, self.size)
        start = max(0, start)
        n = max(0, stop - start)
        if out is None:
            out = np.empty(self._shape(n), dtype=self.dtype)
        flat = out.reshape(-1)
        row_items = self.nmemb
        done = 0
        for i in range(len(self.fsize)):
            a, b = self.foff[i], self.foff[i + 1]
            lo, hi = max(start, a), min(stop, b)
            if lo >= hi:
                continue
            cnt = (hi - lo) * row_items
            with open(os.path.join(self.path, f"{i:06X}"), "rb") as f:
                f.seek(int(lo - a) * row_items * self.dtype.itemsize)
                got = f.readinto(memoryview(flat[done:done + cnt]).cast("B"))
            if got != cnt * self.dtype.itemsize:
                raise IOError(f"{self.path}/{i:06X}: short read")
            done += cnt
        return out

    def __getitem__(self, sl):
        if not isinstance(sl, slice) or sl.step not in (None, 1):
            raise TypeError("blocks support contiguous row slices only")
        start, stop, _ = sl.indices(self.size)
        return self.read(start, stop)


def write_block(path, array, attrs=None, nfile=1):
    """Write ``array`` ((rows,) or (rows, nmemb), C order) as a block at
    ``path`` split over ``nfile`` data files, with ``attrs``."""
    a = np.ascontiguousarray(array)
    if a.ndim == 1:
        nmemb = 1
    elif a.ndim == 2:
        nmemb = a.shape[1]
    else:
        raise ValueError("blocks hold (rows,) or (rows, nmemb) arrays; flatten first")
    rows = a.shape[0]
    nfile = max(1, int(nfile))
    os.makedirs(path, exist_ok=True)
    bounds = [rows * i // nfile for i in range(nfile + 1)]
    lines = [f"DTYPE: {a.dtype.str}\n", f"NMEMB: {nmemb}\n", f"NFILE: {nfile}\n"]
    for i in range(nfile):
        part = a[bounds[i]:bounds[i + 1]]
        with open(os.path.join(path, f"{i:06X}"), "wb") as f:
            f.write(part.tobytes())
        s, fold = checksums(part.tobytes())
        lines.append(f"{i:06X}: {bounds[i + 1] - bounds[i]} : {s} : {fold}\n")
    with open(os.path.join(path, "header"), "w") as f:
        f.writelines(lines)
    _write_attrs(path, attrs or {})


def savefield(data, filename, pm, nfile=1):
    """Mesh ``data`` (N0, N1, N2) as block ``Field`` of ``filename``, flattened
    in C order, with the attributes ``ndarray.shape``, ``BoxSize`` and
    ``Nmesh`` (maputils.py:26-41).  ``pm`` needs ``Nmesh`` and ``BoxSize``."""
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    nmesh = np.asarray(pm.Nmesh, dtype=np.int64).reshape(-1)
    box = np.asarray(pm.BoxSize, dtype=np.float64).reshape(-1)
    write_block(os.path.join(filename, "Field"), np.asarray(data).reshape(-1), nfile=nfile,
                attrs={"ndarray.shape": nmesh, "BoxSize": box, "Nmesh": nmesh})


def loadfield(path, pm=None, local=False):
    """(mesh, Nmesh, BoxSize) from block ``Field`` of ``path``
    (maputils.py:44-76): float64 for 8-byte items, float32 otherwise, shaped
    by the ``Nmesh`` attribute.  One process owns the whole mesh, so
    ``local`` selects all of it."""
    b = BigBlock(os.path.join(path, "Field"))
    attrs = {k: np.squeeze(v) for k, v in b.attrs.items()}
    nmesh = np.atleast_1d(attrs["Nmesh"]).astype(np.int64)
    if nmesh.size == 1:
        nmesh = np.repeat(nmesh, 3)
    dtype = np.float64 if b.dtype.itemsize == 8 else np.float32
    out = b.read().astype(dtype, copy=False).reshape(tuple(int(v) for v in nmesh))
    return out, attrs["Nmesh"], attrs["BoxSize"]


def load_positions(path, device="cuda", dtype=None, block="Position", chunk_rows=1 << 22):
    """Particle positions of block ``block`` under ``path`` (the FastPM
    ``File(path)['Position']`` of runJERALD.py:47-62) as a (Np, 3) torch
    tensor on ``device``.  Rows stream from disk through a pinned host
    buffer of ``chunk_rows`` rows, copied asynchronously while the next chunk
    is read."""
    import torch

    b = BigBlock(os.path.join(path, block))
    if b.nmemb != 3:
        raise ValueError(f"{block} has NMEMB {b.nmemb}, expected 3")
    if b.dtype.kind != "f" or b.dtype.byteorder == ">":
        raise ValueError(f"{block} has DTYPE {b.dtype.str}, expected little-endian float")
    src_t = torch.float32 if b.dtype.itemsize == 4 else torch.float64
    dst_t = dtype or src_t
    dev = torch.device(device)
    out = torch.empty((b.size, 3), dtype=dst_t, device=dev)
    if dev.type != "cuda":
        out.copy_(torch.from_numpy(b.read()))
        return out
    rows = max(1, min(int(chunk_rows), b.size))
    stage = [torch.empty((rows, 3), dtype=src_t, pin_memory=True) for _ in range(2)]
    done = [None, None]
    for k, start in enumerate(range(0, b.size, rows)):
        stop = min(b.size, start + rows)
        s = k & 1
        if done[s] is not None:
            done[s].synchronize()  # the copy out of this staging buffer has finished
        b.read(start, stop, out=stage[s][: stop - start].numpy())
        out[start:stop].copy_(stage[s][: stop - start], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        done[s] = ev
    torch.cuda.synchronize(dev)
    return out
