"""Restart and coordinate files through the library's native readers and writers.

* ``write_restart`` / ``read_restart``: dump_global_velocity_id_mpi / read_dump_global_id_v
  (Source/mod_datadump.f90:113-151, 212-252), ids widened past 99,999 instead of *****;
* ``read_gro`` / ``write_gro``: read_in_Grofile(_box) and the coordinate writer of
  template_generator (Source/mod_GroFile.f90:7-68, 89-159), parsed by column;
* ``system_from_gro``: a glycine/water system (residues TTA and SOL, the reference's input
  files) from a GRO file, with the topology of ``system.assemble``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import system as S
from ._lib import check, lib, ptr

f32, i32 = np.float32, np.int32


def _p(path) -> bytes:
    return str(path).encode()


def write_restart(path, box: float, temp: float, num: int, timeclock: float, global_id, pos, v):
    """One restart file; pos and v (n, 3) and global_id (n,) in any matching order."""
    gid = np.ascontiguousarray(global_id, i32)
    x = np.ascontiguousarray(pos, f32)
    u = np.ascontiguousarray(v, f32)
    if x.shape != (len(gid), 3) or u.shape != x.shape:
        raise ValueError("pos and v must be (n, 3) with one global id per atom")
    check(lib().sofia_write_restart(_p(path), float(box), float(temp), int(num), float(timeclock),
                                    len(gid), ptr(gid), ptr(x), ptr(u)))


def read_restart(path, np_atoms: int):
    """(box, temp, num, timeclock, pos, v) with pos, v (np_atoms, 3) in global-id order."""
    x = np.zeros((np_atoms, 3), f32)
    u = np.zeros((np_atoms, 3), f32)
    box, temp, tc = C.c_float(), C.c_float(), C.c_float()
    num = C.c_int32()
    check(lib().sofia_read_restart(_p(path), int(np_atoms), C.byref(box), C.byref(temp),
                                   C.byref(num), C.byref(tc), ptr(x), ptr(u)))
    return box.value, temp.value, num.value, tc.value, x, u


def read_gro(path):
    """(pos (n, 3) in A, cubic box in A) of a GRO file, atoms in file order."""
    n = C.c_int32()
    box = C.c_float()
    check(lib().sofia_read_gro(_p(path), 0, C.byref(n), None, C.byref(box)))
    x = np.zeros((n.value, 3), f32)
    check(lib().sofia_read_gro(_p(path), n.value, C.byref(n), ptr(x), C.byref(box)))
    return x, box.value


def _names(names, n):
    out = np.zeros((n, 5), np.uint8) + ord(" ")
    for k, s in enumerate(names):
        b = s.encode()[:5]
        out[k, :len(b)] = np.frombuffer(b, np.uint8)
    return out.tobytes()


def write_gro(path, pos, box: float, resid, resname, atomname, title: str = "SOFIA"):
    """GRO file: pos (n, 3) in A, per-atom residue number, residue and atom names."""
    x = np.ascontiguousarray(pos, f32)
    rid = np.ascontiguousarray(resid, i32)
    n = len(x)
    check(lib().sofia_write_gro(_p(path), title.encode(), n, ptr(rid), _names(resname, n),
                                _names(atomname, n), ptr(x), float(box)))


def system_atom_names(s: S.SofiaSystem):
    """(resid, resname, atomname) per atom of a glycine/water system (names of the
    reference's GRO inputs: TTA C1 C2 N O1 O2 H1-H5, SOL OW HW HW)."""
    gly = ["C1", "C2", "N", "O1", "O2", "H1", "H2", "H3", "H4", "H5"]
    resid = np.concatenate([np.repeat(np.arange(1, s.n_gly + 1), 10),
                            np.repeat(np.arange(s.n_gly + 1, s.n_gly + s.n_wat + 1), 3)]).astype(i32)
    resname = ["TTA"] * (10 * s.n_gly) + ["SOL"] * (3 * s.n_wat)
    atomname = gly * s.n_gly + ["OW", "HW", "HW"] * s.n_wat
    return resid, resname, atomname


def system_from_gro(path, seed: int = 20261202, params: S.Params | None = None) -> S.SofiaSystem:
    """Glycine/water system from a GRO file whose residues are TTA (10 atoms in the
    glycine template order) followed by SOL (O, H, H), as the reference's inputs are."""
    pos, box = read_gro(path)
    resnames = []
    with open(path) as f:
        f.readline()
        n = int(f.readline())
        for _ in range(n):
            resnames.append(f.readline()[5:10].strip())
    n_gly_atoms = sum(1 for r in resnames if r != "SOL")
    n_wat_atoms = len(resnames) - n_gly_atoms
    if n_gly_atoms % 10 or n_wat_atoms % 3 or any(r == "SOL" for r in resnames[:n_gly_atoms]):
        raise ValueError("expected 10-atom glycines followed by 3-atom waters")
    params = params or S.Params()
    box32 = f32(box)
    x = np.mod(pos, box32).astype(f32)
    x = np.where(x <= 0, x + box32, x).astype(f32)
    return S.assemble(x, box32, n_gly_atoms // 10, n_wat_atoms // 3, seed, params,
                      np.random.default_rng(seed))
