"""FCIDUMP_chol input: the reference's Hamiltonian file.

The reference writes `FCIDUMP_chol` with h5py (pyscf_interface.py:400-420):
datasets `header` = [nelec, norb, ms, nchol] (int64), `hcore` (N*N), `hcore_mod`
(N*N), `chol` (G*N*N), `energy_core` (scalar) and optionally
`mo_coeffs_up`/`mo_coeffs_dn`; it reads them back in launch_script.py:60-94
with nelec_sp = ((nelec+|ms|)//2, (nelec-|ms|)//2).

h5py is not a dependency here, so `read_hdf5_datasets` reads the subset of the
HDF5 format such files use: superblock version 0 or 1, a root group stored as
a symbol table (version-1 B-tree + local heap), version-1 object headers, and
contiguous or compact datasets of little-endian IEEE floats or integers.
Anything else (chunked/filtered storage, newer superblocks) raises.  The same
data can be kept as `.npz` with the same keys (`save_npz` / `load`).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


class HDF5FormatError(ValueError):
    pass


class _Reader:
    def __init__(self, buf: bytes):
        self.b = buf
        if buf[:8] != _SIG:
            raise HDF5FormatError("not an HDF5 file (signature)")
        ver = buf[8]
        if ver not in (0, 1):
            raise HDF5FormatError(f"superblock version {ver} is not supported (versions 0 and 1 are)")
        self.so, self.sl = buf[13], buf[14]
        if self.so != 8 or self.sl != 8:
            raise HDF5FormatError("only 8-byte offsets and lengths are supported")
        p = 24 + (4 if ver == 1 else 0)  # v1 adds indexed-storage K + reserved
        self.base = self.u(p, 8)
        p += 4 * 8  # base, free-space, EOF, driver-info addresses
        # root group symbol table entry
        self.root_header = self.u(p + 8, 8)
        cache = self.u(p + 16, 4)
        if cache != 1:
            raise HDF5FormatError("root group without a cached symbol table is not supported")
        self.root_btree = self.u(p + 24, 8)
        self.root_heap = self.u(p + 32, 8)

    def u(self, off: int, n: int) -> int:
        return int.from_bytes(self.b[off : off + n], "little")

    def heap_data(self, addr: int) -> int:
        if self.b[addr : addr + 4] != b"HEAP":
            raise HDF5FormatError("local heap signature")
        return self.u(addr + 24, 8)

    def name(self, heap_data: int, off: int) -> str:
        end = self.b.index(b"\x00", heap_data + off)
        return self.b[heap_data + off : end].decode()

    def symbols(self, btree: int, heap: int) -> Dict[str, int]:
        hd = self.heap_data(heap)
        out: Dict[str, int] = {}

        def walk(node: int):
            if self.b[node : node + 4] != b"TREE":
                raise HDF5FormatError("B-tree signature")
            ntype, level, used = self.b[node + 4], self.b[node + 5], self.u(node + 6, 2)
            if ntype != 0:
                raise HDF5FormatError("expected a group B-tree")
            p = node + 8 + 16  # siblings
            children = []
            for _ in range(used):
                p += 8  # key
                children.append(self.u(p, 8))
                p += 8
            for c in children:
                if level > 0:
                    walk(c)
                    continue
                if self.b[c : c + 4] != b"SNOD":
                    raise HDF5FormatError("symbol table node signature")
                n = self.u(c + 6, 2)
                e = c + 8
                for _ in range(n):
                    out[self.name(hd, self.u(e, 8))] = self.u(e + 8, 8)
                    e += 40

        walk(btree)
        return out

    def messages(self, addr: int):
        if self.b[addr] != 1:
            raise HDF5FormatError(f"object header version {self.b[addr]} is not supported")
        total = self.u(addr + 2, 2)
        blocks = [(addr + 16, self.u(addr + 8, 4))]
        seen = 0
        while blocks and seen < total:
            p, size = blocks.pop(0)
            end = p + size
            while p + 8 <= end and seen < total:
                mtype, msize = self.u(p, 2), self.u(p + 2, 2)
                data = p + 8
                seen += 1
                if mtype == 0x0010:  # continuation
                    blocks.append((self.u(data, 8), self.u(data + 8, 8)))
                else:
                    yield mtype, data, msize
                p = data + msize

    def dataset(self, addr: int) -> np.ndarray:
        shape = dtype = layout = None
        for mtype, d, size in self.messages(addr):
            if mtype == 0x0001:
                shape = self._dataspace(d)
            elif mtype == 0x0003:
                dtype = self._datatype(d)
            elif mtype == 0x0008:
                layout = self._layout(d)
            elif mtype == 0x000B:
                raise HDF5FormatError("filtered (compressed) datasets are not supported")
        if shape is None or dtype is None or layout is None:
            raise HDF5FormatError("dataset header lacks dataspace, datatype or layout")
        kind, a, n = layout
        count = int(np.prod(shape)) if shape else 1
        nbytes = count * dtype.itemsize
        if kind == "contiguous":
            if a == (1 << 64) - 1:
                raw = b"\x00" * nbytes  # never written: fill value zero
            else:
                raw = self.b[self.base + a : self.base + a + nbytes]
        else:
            raw = self.b[a : a + nbytes]
        if len(raw) != nbytes:
            raise HDF5FormatError("dataset storage is truncated")
        arr = np.frombuffer(raw, dtype=dtype, count=count).copy()
        return arr.reshape(shape) if shape else arr.reshape(())

    def _dataspace(self, d: int):
        ver, rank, flags = self.b[d], self.b[d + 1], self.b[d + 2]
        if ver == 1:
            p = d + 8
        elif ver == 2:
            p = d + 4
            if self.b[d + 3] == 2:
                return ()
        else:
            raise HDF5FormatError(f"dataspace version {ver}")
        return tuple(self.u(p + 8 * i, 8) for i in range(rank))

    def _datatype(self, d: int) -> np.dtype:
        cls = self.b[d] & 0x0F
        bits = self.b[d + 1]
        size = self.u(d + 4, 4)
        if bits & 0x01:
            raise HDF5FormatError("big-endian data is not supported")
        if cls == 0:
            return np.dtype(("<i" if bits & 0x08 else "<u") + str(size))
        if cls == 1 and size in (4, 8) and not bits & 0x40:
            return np.dtype("<f" + str(size))
        raise HDF5FormatError(f"datatype class {cls} size {size} is not supported")

    def _layout(self, d: int):
        ver = self.b[d]
        if ver == 3:
            cls = self.b[d + 1]
            if cls == 0:
                return ("compact", d + 4, self.u(d + 2, 2))
            if cls == 1:
                return ("contiguous", self.u(d + 2, 8), self.u(d + 10, 8))
            raise HDF5FormatError("chunked datasets are not supported")
        if ver in (1, 2):
            rank, cls = self.b[d + 1], self.b[d + 2]
            if cls == 1:
                return ("contiguous", self.u(d + 8, 8), None)
            if cls == 0:
                p = d + 8 + 4 * rank
                return ("compact", p + 4, self.u(p, 4))
            raise HDF5FormatError("chunked datasets are not supported")
        raise HDF5FormatError(f"layout message version {ver}")


def read_hdf5_datasets(path: str) -> Dict[str, np.ndarray]:
    """Every dataset in the root group of an HDF5 file (subset described above)."""
    with open(path, "rb") as f:
        buf = f.read()
    r = _Reader(buf)
    return {k: r.dataset(a) for k, a in r.symbols(r.root_btree, r.root_heap).items()}


@dataclass
class Hamiltonian:
    h0: float
    h1: np.ndarray  # (N, N)
    chol: np.ndarray  # (G, N, N)
    norb: int
    nelec_sp: Tuple[int, int]
    h1_mod: Optional[np.ndarray] = None
    mo_coeffs: Optional[Tuple[np.ndarray, np.ndarray]] = None


def _from_datasets(d: Dict[str, np.ndarray]) -> Hamiltonian:
    nelec, norb, ms, nchol = (int(x) for x in np.asarray(d["header"]).reshape(-1)[:4])
    h0 = float(np.asarray(d["energy_core"]).reshape(()))
    h1 = np.asarray(d["hcore"], dtype=np.float64).reshape(norb, norb)
    chol = np.asarray(d["chol"], dtype=np.float64).reshape(-1, norb, norb)
    if chol.shape[0] != nchol:
        raise ValueError(f"header says {nchol} Cholesky vectors, file holds {chol.shape[0]}")
    hmod = np.asarray(d["hcore_mod"], dtype=np.float64).reshape(norb, norb) if "hcore_mod" in d else None
    mo = None
    if "mo_coeffs_up" in d and "mo_coeffs_dn" in d:
        mo = (np.asarray(d["mo_coeffs_up"]), np.asarray(d["mo_coeffs_dn"]))
    nelec_sp = ((nelec + abs(ms)) // 2, (nelec - abs(ms)) // 2)
    return Hamiltonian(h0, h1, chol, norb, nelec_sp, hmod, mo)


def load(path: str) -> Hamiltonian:
    """FCIDUMP_chol (HDF5) or its .npz copy -> Hamiltonian (launch_script.read_fcidump)."""
    if path.endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            return _from_datasets({k: z[k] for k in z.files})
    return _from_datasets(read_hdf5_datasets(path))


def save_npz(path: str, h0: float, h1: np.ndarray, chol: np.ndarray, nelec: Tuple[int, int],
             h1_mod: Optional[np.ndarray] = None) -> None:
    """Writes the FCIDUMP_chol datasets (same names and flattening) to an .npz."""
    norb = h1.shape[-1]
    chol = np.asarray(chol).reshape(-1, norb * norb)
    if h1_mod is None:
        L = chol.reshape(-1, norb, norb)
        h1_mod = h1 - 0.5 * np.einsum("gik,gjk->ij", L, L)  # pyscf_interface.py:99-100
    ms = nelec[0] - nelec[1]
    np.savez(path, header=np.array([nelec[0] + nelec[1], norb, ms, chol.shape[0]], dtype=np.int64),
             hcore=np.asarray(h1).reshape(-1), hcore_mod=np.asarray(h1_mod).reshape(-1), chol=chol.reshape(-1),
             energy_core=np.float64(h0))


def ham_data(h: Hamiltonian) -> dict:
    """The reference's ham_data dict (launch_script.py:213-243): h1 duplicated per spin."""
    N = h.norb
    return {"h0": h.h0, "h1": np.array([h.h1, h.h1]), "chol": h.chol.reshape(-1, N * N), "ene0": 0.0}
