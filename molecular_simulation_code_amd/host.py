"""Host side of the MI355X hot path, mirroring the reference's interface.

``SofiaGPU`` holds the state the reference keeps in ``module global``
(Source/global.f90) for one system and exposes the hot-path subroutines under their
reference names -- ``neighbor_wrapper``, ``check_neighbor``,
``lj_cut_coul_dsf_nonewton``, ``lj_cut_coul_long_nonewton``, ``correct_12_bonds`` ...,
``nve_v``, ``nve_x``, ``pbc``, ``shake``, ``langevin`` -- each a thin call through the
C ABI of libsofia_hip.so.  Like the Fortran routines they read and update the object's
arrays (position, v, ff, energies) in the reference's layout and precision, including the
permanent cell-sort permutation of neighbor_wrapper.

``Engine`` is the device-resident path: one or many replicas advanced in fused MD segments
in FP64 or mixed precision, with the state kept on the GPU between calls.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import system as S
from ._lib import (COUL_DSF, COUL_EWALD, NENERGY, PREC_FP64, PREC_MIXED, SofiaParams,
                   SofiaThermostat, check, lib, ptr)

f32, f64, i32 = np.float32, np.float64, np.int32


def make_params(d: dict, precision: int = PREC_FP64, coul_flag: int = COUL_DSF,
                params: S.Params | None = None, neigh_alloc0: int = S.NEIGH_ALLOC0) -> SofiaParams:
    """sofia_params from the derived constants of ``system.Params.derived``."""
    p = params or S.Params()
    return SofiaParams(ntypes=S.NTYPES, rcut=float(d["rcut"]), cut_coul=float(d["cut_coul"]),
                       pad=float(d["pad"]), coulpre=float(S.COULPRE), dt=float(d["dt"]),
                       kcal2amu=float(S.KCAL2AMU), rcut2=float(d["rcut2"]),
                       cut_coulsq=float(d["cut_coulsq"]), rv2=float(d["rv2"]),
                       rvrc2=float(d["rvrc2"]), alpha=float(d["alpha"]), gewald=float(d["gewald"]),
                       e_shift=float(d["e_shift"]), f_shift=float(d["f_shift"]),
                       ewald_p=float(S.EWALD_P), ewald_f=float(S.EWALD_F), a1=float(S.A1),
                       a2=float(S.A2), a3=float(S.A3), a4=float(S.A4), a5=float(S.A5),
                       my_pis=float(S.MY_PIS), my_pis_inv=float(S.MY_PIS_INV),
                       amu2e=float(S.AMU2E), fudge_12=float(p.fudge_12), fudge_13=float(p.fudge_13),
                       fudge_14_coul=float(p.fudge_14_coul), shake_tol=float(S.SHAKE_TOL),
                       shake_maxiter=S.SHAKE_MAXITER, precision=precision, coul_flag=coul_flag,
                       neigh_alloc0=neigh_alloc0)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _forcefield(ctx):
    lj = S.lj_tables()
    arrs = [_c(x, f32) for x in lj[:4]] + [_c(x, f64) for x in lj[4:]] + [_c(S.type_mass(), f64)]
    check(lib().sofia_set_forcefield(ctx, *[ptr(a) for a in arrs]))
    return arrs


def torsion_list(s: S.SofiaSystem) -> np.ndarray:
    """Torsions in geometric order a-b-c-d (1-based), one per topology dihedral."""
    rows = []
    base = 0
    for _ in range(s.n_gly):
        for a, b, c, d, t in S.GLYCINE.diheds:
            rows.append((base + a, base + b, base + c, base + d, t))
        base += 10
    return np.array(rows, i32).reshape(-1, 5)


def _bonded_coeffs(ctx, tors: np.ndarray, ntor_rep=None):
    bk = _c([c[0] for c in S.BOND_COEFF], f64); br = _c([c[1] for c in S.BOND_COEFF], f64)
    ak = _c([c[0] for c in S.ANGLE_COEFF], f64)
    ath = _c(np.radians([c[1] for c in S.ANGLE_COEFF]), f64)
    tk = _c([c[0] for c in S.DIHED_COEFF], f64); tn = _c([c[1] for c in S.DIHED_COEFF], f64)
    ts = _c([c[2] for c in S.DIHED_COEFF], f64)
    tors = _c(tors, i32)
    keep = [bk, br, ak, ath, tk, tn, ts, tors]
    nr = None if ntor_rep is None else _c(ntor_rep, i32)
    check(lib().sofia_set_bonded_coeffs(ctx, len(bk), ptr(bk), ptr(br), len(ak), ptr(ak), ptr(ath),
                                        len(tk), ptr(tk), ptr(tn), ptr(ts), ptr(tors), len(tors),
                                        ptr(nr)))
    return keep


def _impropers(ctx, lists, coul_flag=COUL_DSF):
    """Improper items (i1, i2, i3 central, i4, type) per replica, in the form the reference's
    force wrapper of that electrostatics calls: improper_harmonic_lammps14 with DSF
    (Source/mod_force.f90:542-544), the cvff impro_amber_force with Ewald
    (Source/mod_force.f90:141-143) and the topology's (K, d, n)."""
    il = _c(np.concatenate(lists) if lists else np.zeros((0, 5), i32), i32)
    nr = _c([len(x) for x in lists], i32)
    if coul_flag == COUL_EWALD:
        ik = _c([c[0] for c in S.IMPRO_COEFF], f64)
        isg = _c([c[1] for c in S.IMPRO_COEFF], i32)
        imt = _c([c[2] for c in S.IMPRO_COEFF], i32)
        check(lib().sofia_set_impropers_cvff(ctx, len(ik), ptr(ik), ptr(isg), ptr(imt), ptr(il),
                                             len(il), ptr(nr)))
        return [ik, isg, imt, il, nr]
    ik = _c([c[0] for c in S.IMPRO_K_CHI0], f64)
    ic = _c([c[1] for c in S.IMPRO_K_CHI0], f64)
    check(lib().sofia_set_impropers(ctx, len(ik), ptr(ik), ptr(ic), ptr(il), len(il), ptr(nr)))
    return [ik, ic, il, nr]


class SofiaGPU:
    """One system with the reference's module-global state and subroutine names."""

    def __init__(self, system: S.SofiaSystem, precision: int = PREC_FP64):
        self.sys = system
        self.d = system.derived()
        self.np = system.np
        self.params = make_params(self.d, precision, params=system.params)
        h = C.c_void_p()
        check(lib().sofia_ctx_create(C.byref(self.params), self.np, float(self.d["box"]), C.byref(h)))
        self.ctx = h
        self._keep = _forcefield(h)
        # module-global state (reference layout and precision)
        self.position = system.pos.astype(f32).copy()
        self.type = system.types.astype(i32).copy()
        self.q = system.q.astype(f32).copy()
        self.v = system.v.astype(f32).copy()
        self.global_id = system.global_id.astype(i32).copy()
        self.mol = system.mol.astype(i32).copy()
        self.num1bond = system.num1bond.copy(); self.num2bond = system.num2bond.copy()
        self.num3bond = system.num3bond.copy(); self.specbond = system.specbond.copy()
        self.bondlist = system.bondlist.copy(); self.anglelist = system.anglelist.copy()
        self.dihedlist = system.dihedlist.copy()
        self.shake_num = system.shake_num.copy(); self.shake_atoms = system.shake_atoms.copy()
        self.shake_bond = system.shake_bond.copy()
        self.ff = np.zeros((self.np, 3), f32)
        self.drx = np.zeros(self.np, f32); self.dry = np.zeros(self.np, f32); self.drz = np.zeros(self.np, f32)
        self.potential = 0.0; self.e_coul = 0.0; self.e_coul_14 = 0.0; self.pot_14 = 0.0
        self.e_coul_bond = 0.0; self.e_coul_angle = 0.0; self.e_coul_dihed = 0.0
        self.virial = np.zeros(3); self.ke = 0.0
        self.neigh_alloc = S.NEIGH_ALLOC0
        self.nn_builds = 0
        self._upload_all()
        # bond / angle / torsion / improper coefficients for bonded_forces()
        self._keep += _bonded_coeffs(h, torsion_list(system))
        self._keep += _impropers(h, [system.improlist])
        self.e_bond = self.e_angle = self.e_dihed = self.e_improper = 0.0

    # ------------------------------------------------------------------ plumbing
    def close(self):
        if getattr(self, "ctx", None):
            lib().sofia_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ainfo(self):
        a = np.empty((self.np, 4), f32)
        a[:, :3] = self.position
        a[:, 3] = self.type.view(f32)
        return a

    def _upload_all(self):
        a = self.ainfo()
        check(lib().sofia_put_state(self.ctx, ptr(a), ptr(_c(self.q, f32)), ptr(_c(self.v, f32)),
                                    ptr(_c(self.global_id, i32))))
        check(lib().sofia_set_exclusions(self.ctx, ptr(_c(self.num1bond, i32)), ptr(_c(self.num2bond, i32)),
                                         ptr(_c(self.num3bond, i32)), ptr(_c(self.specbond, i32))))
        check(lib().sofia_set_bonded(self.ctx, ptr(_c(self.bondlist, i32)), len(self.bondlist), None,
                                     ptr(_c(self.anglelist, i32)), len(self.anglelist), None,
                                     ptr(_c(self.dihedlist, i32)), len(self.dihedlist), None))
        check(lib().sofia_set_shake(self.ctx, len(self.shake_num), None, ptr(_c(self.shake_num, i32)),
                                    ptr(_c(self.shake_atoms, i32)), ptr(_c(self.shake_bond, f64))))
        self._put_dr()

    def _put_pos(self):
        a = self.ainfo()
        check(lib().sofia_put_state(self.ctx, ptr(a), None, ptr(_c(self.v, f32)), None))

    def _get_pos(self):
        a = np.empty((self.np, 4), f32)
        v = np.empty((self.np, 3), f32)
        gid = np.empty(self.np, i32)
        check(lib().sofia_get_state(self.ctx, ptr(a), ptr(v), ptr(gid)))
        self.position = a[:, :3].copy()
        self.type = a[:, 3].view(i32).copy()
        self.v = v
        self.global_id = gid

    def _put_dr(self):
        check(lib().sofia_put_displacement(self.ctx, ptr(_c(self.drx, f32)), ptr(_c(self.dry, f32)),
                                           ptr(_c(self.drz, f32))))

    def _get_dr(self):
        check(lib().sofia_get_displacement(self.ctx, ptr(self.drx), ptr(self.dry), ptr(self.drz)))

    # ------------------------------------------------------------------ reference routines
    def neighbor_wrapper(self, step: int = 0) -> np.ndarray:
        """Source/mod_neighbor.f90:8-86.  Returns tt (old -> new, 1-based) and re-indexes
        every host array exactly as sort_posit_charge/sort_* do."""
        self._put_pos()
        tt = np.empty(self.np, i32)
        check(lib().sofia_neighbor_wrapper(self.ctx, step, ptr(tt)))
        self._get_pos()
        self.permute_host(tt)
        return tt

    def permute_host(self, tt: np.ndarray):
        """Apply the sort permutation tt (old -> new, 1-based) to the host's own per-atom arrays
        and atom ids, as sort_posit_charge and the sort_* routines do
        (Source/mod_neighbor.f90:946-1168); position, v and global_id come back from the device
        already sorted.  Clears drx/dry/drz and reads the list stride."""
        self.nn_builds += 1
        inv = np.empty(self.np, np.intp)              # new position -> old index
        inv[tt - 1] = np.arange(self.np)
        t0 = np.zeros(self.np + 1, i32)               # 1-based id -> new id; 0 (empty) stays 0
        t0[1:] = tt
        self.q = self.q[inv]
        self.mol = self.mol[inv]
        self.num1bond = self.num1bond[inv]; self.num2bond = self.num2bond[inv]
        self.num3bond = self.num3bond[inv]
        self.specbond = t0[self.specbond[inv]]
        self.bondlist[:, :2] = t0[self.bondlist[:, :2]]
        self.anglelist[:, :3] = t0[self.anglelist[:, :3]]
        self.dihedlist[:, :4] = t0[self.dihedlist[:, :4]]
        self.shake_atoms = t0[self.shake_atoms]
        self.drx[:] = 0; self.dry[:] = 0; self.drz[:] = 0
        na = C.c_int32()
        check(lib().sofia_get_nlist(self.ctx, None, None, C.byref(na)))
        self.neigh_alloc = na.value

    def get_nlist(self):
        na = C.c_int32()
        check(lib().sofia_get_nlist(self.ctx, None, None, C.byref(na)))
        nl = np.empty(self.np * na.value, i32)
        nn = np.empty(self.np, i32)
        check(lib().sofia_get_nlist(self.ctx, ptr(nl), ptr(nn), C.byref(na)))
        return nl, nn, na.value

    def get_cells(self):
        from . import system as _s
        rn, nT, nD = _s.cell_geometry(float(self.d["box"]), float(self.d["rcut"]), float(self.d["cut_coul"]))
        st = np.empty(nT, i32); en = np.empty(nT, i32); ab = np.empty(self.np, i32)
        a, b, r = C.c_int32(), C.c_int32(), C.c_float()
        check(lib().sofia_get_cells(self.ctx, ptr(st), ptr(en), ptr(ab), C.byref(a), C.byref(b), C.byref(r)))
        return st, en, ab, (np.float32(r.value), a.value, b.value)

    def check_neighbor(self, step: int = 0) -> int:
        """Source/mod_neighbor.f90:895-913 on drx/dry/drz."""
        self._put_dr()
        f = C.c_int32()
        check(lib().sofia_check_neighbor(self.ctx, C.byref(f)))
        return f.value

    def _nonbond(self, fn, with64=True):
        self._put_pos()
        ff = np.empty((self.np, 3), f32)
        ff64 = np.empty((self.np, 3), f64) if with64 else None
        pot, ec = C.c_double(0.0), C.c_double(0.0)
        vir = np.zeros(3, f64)
        if with64:
            check(fn(self.ctx, ptr(ff), ptr(ff64), C.byref(pot), C.byref(ec), ptr(vir)))
        else:
            check(fn(self.ctx, ptr(ff), C.byref(pot), C.byref(ec), ptr(vir)))
        self.ff = ff
        self.ff64 = ff64
        self.potential += pot.value
        self.e_coul += ec.value
        self.virial = vir
        return ff

    def lj_cut_coul_dsf_nonewton(self, step: int = 0):
        """lj_cut_coul_dsf_nonewton_NOMIC, Source/mod_force_nonbond.f90:105-196 (FP64)."""
        return self._nonbond(lib().sofia_lj_cut_coul_dsf)

    def lj_cut_coul_dsf_nonewton_mixed(self, step: int = 0):
        """FP32 pair arithmetic with int64 fixed-point force sums."""
        return self._nonbond(lib().sofia_lj_cut_coul_dsf_mixed)

    def lj_cut_coul_dsf_nonewton_async(self, mixed: bool = False):
        """Start the DSF pair kernel and return at once (the reference's offload ... signal);
        nonbond_wait() delivers ff, potential, e_coul and the virial."""
        self._put_pos()
        check(lib().sofia_lj_cut_coul_dsf_async(self.ctx, 1 if mixed else 0))

    def nonbond_wait(self):
        """Wait for the pair kernel started by lj_cut_coul_dsf_nonewton_async (offload_wait)."""
        ff = np.empty((self.np, 3), f32)
        pe, ec = C.c_double(0.0), C.c_double(0.0)
        vir = np.zeros(3, f64)
        check(lib().sofia_nonbond_wait(self.ctx, ptr(ff), C.byref(pe), C.byref(ec), ptr(vir)))
        self.ff = ff
        self.potential += pe.value
        self.e_coul += ec.value
        self.virial = vir
        return ff

    def set_tuning(self, lanes: int = 16, block: int = 1024, list_waves: int = 4):
        """Pair-kernel lanes per atom (0: the compacting mixed kernel), block, list waves."""
        check(lib().sofia_set_tuning(self.ctx, lanes, block, list_waves))

    def lj_cut_coul_long_nonewton(self, step: int = 0):
        """lj_cut_coul_long_nonewton_NOMIC, Source/mod_force_nonbond.f90:300-400."""
        return self._nonbond(lib().sofia_lj_cut_coul_long, with64=False)

    def _correct(self, which):
        self._put_pos()
        ff = np.empty((self.np, 3), f32)
        ec = C.c_double(0.0)
        vir = np.zeros(3, f64)
        check(lib().sofia_correct_bonds(self.ctx, which, ptr(ff), C.byref(ec), ptr(vir)))
        return ff, ec.value, vir

    def correct_12_bonds(self):
        """Source/mod_force_correction.f90:6-93 -> (ff_bond, e_coul_bond, v_bond)."""
        ff, e, v = self._correct(12)
        self.ff_bond, self.e_coul_bond, self.v_bond = ff, self.e_coul_bond + e, v
        return ff

    def correct_13_bonds(self):
        """Source/mod_force_correction.f90:95-184."""
        ff, e, v = self._correct(13)
        self.ff_angle, self.e_coul_angle, self.v_angle = ff, self.e_coul_angle + e, v
        return ff

    def correct_14_bonds(self):
        """Source/mod_force_correction.f90:187-278."""
        ff, e, v = self._correct(14)
        self.ff_dihed, self.e_coul_dihed, self.v_dihed = ff, self.e_coul_dihed + e, v
        return ff

    def pair14_dsf(self):
        """1-4 pair of dihedral_opls_DSF, Source/mod_force_dihed.f90:1003-1048 -> ffh."""
        self._put_pos()
        ff = np.empty((self.np, 3), f32)
        e1, e2 = C.c_double(0.0), C.c_double(0.0)
        check(lib().sofia_pair14_dsf(self.ctx, ptr(ff), C.byref(e1), C.byref(e2)))
        self.ffh = ff
        self.e_coul_14 += e1.value
        self.pot_14 += e2.value
        return ff

    def bonded_forces(self):
        """bond_harmonic + angle_harmonic + torsions + improper_harmonic_lammps14 of
        force_wrapper_DSF_NOMIC (Source/mod_force.f90:510-554) -> their summed forces;
        e_bond / e_angle / e_dihed / e_improper are added to."""
        self._put_pos()
        ff = np.empty((self.np, 3), f32)
        e = [C.c_double(0.0) for _ in range(4)]
        check(lib().sofia_bonded_forces(self.ctx, ptr(ff), *[C.byref(x) for x in e]))
        self.ff_bonded = ff
        self.e_bond += e[0].value
        self.e_angle += e[1].value
        self.e_dihed += e[2].value
        self.e_improper += e[3].value
        return ff

    def nve_v(self, dt_partial: float):
        """Source/mod_ensemble.f90:11-40 with the current ff."""
        self._put_pos()
        ke = C.c_double()
        check(lib().sofia_nve_v(self.ctx, float(np.float32(dt_partial)), ptr(_c(self.ff, f32)), C.byref(ke)))
        self._get_pos()
        self.ke = ke.value

    def nve_x(self):
        """Source/mod_ensemble.f90:48-87 (position and drx/dry/drz)."""
        self._put_pos()
        self._put_dr()
        check(lib().sofia_nve_x(self.ctx))
        self._get_pos()
        self._get_dr()

    def berendsen(self, mu: float):
        """Source/mod_ensemble.f90:167-205 with the scale factor mu computed by the host from
        the pressure: positions and drx/dry/drz on the device, then box / ibox (real*4)."""
        self._put_pos()
        self._put_dr()
        m = np.array([mu], f64)
        b = np.zeros(1, f32)
        check(lib().sofia_berendsen(self.ctx, ptr(m), ptr(b)))
        self._get_pos()
        self._get_dr()
        self.d["box"] = f32(b[0])
        self.d["ibox"] = f32(1.0) / f32(b[0])
        return float(b[0])

    def pbc(self, step: int = 0):
        """Source/mod_pbc.f90:6-39."""
        self._put_pos()
        check(lib().sofia_pbc(self.ctx))
        self._get_pos()

    def shake(self, iflag: int = 1):
        """Source/mod_shake.f90:8-85: constraint forces added to ff."""
        self._put_pos()
        ff = _c(self.ff, f32).copy()
        vir = np.zeros(3, f64)
        mit = C.c_int32()
        check(lib().sofia_shake(self.ctx, iflag, ptr(ff), ptr(vir), C.byref(mit)))
        self.ff = ff
        self.virial = self.virial + vir
        self.shake_iters = mit.value
        return ff

    def langevin(self, fric: float, gconst: float, seed: int, step: int):
        """Source/mod_langevin.f90:15-43 with a counter-based generator."""
        self._put_pos()
        out = np.empty((self.np, 3), f32)
        check(lib().sofia_langevin(self.ctx, float(fric), float(gconst), seed, step, ptr(out)))
        self.ff_lang = out
        return out


class PinnedArrays:
    """numpy arrays in page-locked host memory (sofia_host_alloc); release() frees them,
    after which the arrays must not be used."""

    def __init__(self):
        self._bufs = []

    def empty(self, shape, dtype) -> np.ndarray:
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        check(lib().sofia_host_alloc(nbytes, C.byref(p)))
        self._bufs.append(p)
        return np.frombuffer((C.c_char * nbytes).from_address(p.value), dtype=dtype).reshape(shape)

    def release(self):
        for p in self._bufs:
            lib().sofia_host_free(p)
        self._bufs = []


class DropinStep:
    """The reference's DSF time step (Source/mod_integrate.f90:270-308 with
    force_wrapper_DSF_NOMIC, Source/mod_force.f90:510-554) driven through the drop-in entry
    points the way a Fortran host with page-locked arrays would call them: each routine
    uploads the arrays it reads and downloads the arrays it writes, and nothing else.

    * nve_v: v, ff -> v;  nve_x: position, v, drx/dry/drz -> position, drx/dry/drz;
    * pbc: position -> position;  check_neighbor: drx/dry/drz;
    * force_wrapper_DSF_NOMIC (the DSF pair kernel, the 1-4 pairs and the bonded terms,
      merged on the device as the host merges the three routines' outputs): position -> ff;
    * shake(1): position, v, ff -> ff.

    * neighbor_wrapper, when check_neighbor asks for it: position, v -> position, v,
      global_id, then the host permutes its own per-atom arrays and topology.

    The trajectory equals SofiaGPU's step for step."""

    def __init__(self, g: "SofiaGPU"):
        self.g = g
        self.mem = PinnedArrays()
        n = g.np
        self.pos = self.mem.empty((n, 4), f32)       # type(ainfo): x, y, z, type
        self.v = self.mem.empty((n, 3), f32)
        self.ff = self.mem.empty((n, 3), f32)
        self.dr = self.mem.empty((3, n), f32)
        # potential, e_coul, e_coul_14, pot_14, e_bond, e_angle, e_dihed, e_improper (added to)
        self.energies = np.zeros(8, f64)
        # host arrays that sofia_get_topology fills in place after each list build
        g.q = np.ascontiguousarray(g.q, f32)
        for k in ("num1bond", "num2bond", "num3bond", "specbond", "bondlist", "anglelist",
                  "dihedlist", "shake_atoms"):
            setattr(g, k, np.ascontiguousarray(getattr(g, k), i32))
        self.dthalf = float(g.sys.derived()["dthalf"])
        self._from_g()

    def close(self):
        self.pos = self.v = self.ff = self.dr = None
        self.mem.release()

    def _from_g(self):
        g = self.g
        self.pos[:, :3] = g.position
        self.pos[:, 3] = g.type.view(f32)
        self.v[:] = g.v
        self.ff[:] = g.ff
        self.dr[0], self.dr[1], self.dr[2] = g.drx, g.dry, g.drz

    def _to_g(self):
        g = self.g
        g.position = self.pos[:, :3].copy()
        g.type = self.pos[:, 3].copy().view(i32)
        g.v = self.v.copy()
        g.ff = self.ff.copy()
        g.drx, g.dry, g.drz = self.dr[0].copy(), self.dr[1].copy(), self.dr[2].copy()

    def _put(self, pos=False, v=False):
        check(lib().sofia_put_state(self.g.ctx, ptr(self.pos) if pos else None, None,
                                    ptr(self.v) if v else None, None))

    def _get(self, pos=False, v=False):
        check(lib().sofia_get_state(self.g.ctx, ptr(self.pos) if pos else None,
                                    ptr(self.v) if v else None, None))

    def _dr(self, put: bool):
        fn = lib().sofia_put_displacement if put else lib().sofia_get_displacement
        check(fn(self.g.ctx, ptr(self.dr[0]), ptr(self.dr[1]), ptr(self.dr[2])))

    def neighbor_wrapper(self):
        """neighbor_wrapper (Source/mod_neighbor.f90:8-86): position and v in (q and
        global_id are already on the device, which permutes them itself); the device sorts
        the atoms and builds the list; position, v and global_id out; then the host permutes
        its own per-atom arrays and topology."""
        g = self.g
        tt = np.empty(g.np, i32)
        check(lib().sofia_put_state(g.ctx, ptr(self.pos), None, ptr(self.v), None))
        check(lib().sofia_neighbor_wrapper(g.ctx, 0, ptr(tt)))
        check(lib().sofia_get_state(g.ctx, ptr(self.pos), ptr(self.v), ptr(g.global_id)))
        # the arrays the device holds come back permuted; the host permutes only the rest
        check(lib().sofia_get_topology(g.ctx, ptr(g.q), ptr(g.num1bond), ptr(g.num2bond),
                                       ptr(g.num3bond), ptr(g.specbond), ptr(g.bondlist),
                                       ptr(g.anglelist), ptr(g.dihedlist), None,
                                       ptr(g.shake_atoms)))
        inv = np.empty(g.np, np.intp)
        inv[tt - 1] = np.arange(g.np)
        g.mol = g.mol[inv]
        g.nn_builds += 1
        na = C.c_int32()
        check(lib().sofia_get_nlist(g.ctx, None, None, C.byref(na)))
        g.neigh_alloc = na.value
        self.dr[:] = 0

    def export(self):
        """Copy the step's position, type, v, ff and drx/dry/drz into the SofiaGPU host arrays."""
        self._to_g()

    def nve_v(self):
        ke = C.c_double()
        self._put(v=True)
        check(lib().sofia_nve_v(self.g.ctx, float(np.float32(self.dthalf)), ptr(self.ff), C.byref(ke)))
        self._get(v=True)

    def forces(self):
        """force_wrapper_DSF_NOMIC (pair, 1-4 and bonded forces, merged on the device) then
        shake(1)."""
        ctx = self.g.ctx
        vir = np.zeros(3, f64)
        self._put(pos=True)
        check(lib().sofia_force_wrapper_dsf(ctx, ptr(self.ff), ptr(self.energies), ptr(vir)))
        self._put(pos=True, v=True)
        mit = C.c_int32()
        check(lib().sofia_shake(ctx, 1, ptr(self.ff), ptr(vir), C.byref(mit)))

    def step(self) -> bool:
        """One time step; True when the list was rebuilt."""
        ctx = self.g.ctx
        self.nve_v()
        self._put(pos=True, v=True)
        self._dr(True)
        check(lib().sofia_nve_x(ctx))
        self._get(pos=True)
        self._dr(False)
        self._put(pos=True)
        check(lib().sofia_pbc(ctx))
        self._get(pos=True)
        self._dr(True)
        flag = C.c_int32()
        check(lib().sofia_check_neighbor(ctx, C.byref(flag)))
        rebuilt = bool(flag.value)
        if rebuilt:
            self.neighbor_wrapper()
        self.forces()
        self.nve_v()
        return rebuilt


_TUNE_CACHE: dict = {}
# threads per workgroup of the mixed pair kernel: each workgroup stages the DSF table in LDS,
# so 1024-thread workgroups load it a quarter as often as 256-thread ones (-4.8 % per launch)
PAIR_BLOCK = 1024


def _tune_cache_file():
    import os
    return os.environ.get("SOFIA_TUNE_CACHE")


def _tune_cache_get(key: str):
    if key in _TUNE_CACHE:
        return _TUNE_CACHE[key]
    path = _tune_cache_file()
    if path:
        import json
        try:
            with open(path) as f:
                v = json.load(f).get(key)
        except (OSError, ValueError):
            v = None
        if v is not None:
            _TUNE_CACHE[key] = tuple(v)
            return _TUNE_CACHE[key]
    return None


def _tune_cache_put(key: str, value: tuple):
    _TUNE_CACHE[key] = value
    path = _tune_cache_file()
    if path:
        import json
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        data[key] = list(value)
        with open(path, "w") as f:
            json.dump(data, f, indent=1)


@dataclass
class ReplicaSpec:
    system: S.SofiaSystem
    temp: float = 300.0
    gamma_ps: float = 0.0
    seed: int = 1


class Engine:
    """Device-resident MD for one or many replicas (a single system is a batch of one)."""

    def __init__(self, replicas: list, precision: int = PREC_MIXED, tuning=None,
                 with_bonded: bool = True, coul_flag: int = COUL_DSF):
        if not replicas:
            raise ValueError("no replicas")
        reps = [r if isinstance(r, ReplicaSpec) else ReplicaSpec(r) for r in replicas]
        self.reps = reps
        s0 = reps[0].system
        self.d = s0.derived()
        self.params = make_params(self.d, precision, coul_flag, params=s0.params)
        n = np.array([r.system.np for r in reps], i32)
        boxes = np.array([float(r.system.box) for r in reps], f32)
        self.natoms = n
        self.N = int(n.sum())
        h = C.c_void_p()
        check(lib().sofia_ctx_create_batch(C.byref(self.params), len(reps), ptr(n), ptr(boxes), C.byref(h)))
        self.ctx = h
        self._keep = _forcefield(h)
        # every array handed to the C ABI is bound to a name for the duration of the call
        def cat(f, dtype):
            return np.ascontiguousarray(np.concatenate([f(r.system) for r in reps]), dtype=dtype)

        a = np.empty((self.N, 4), f32)
        a[:, :3] = cat(lambda s: s.pos, f32)
        a[:, 3] = cat(lambda s: s.types, i32).view(f32)
        q = cat(lambda s: s.q, f32)
        v = cat(lambda s: s.v, f32)
        gid = cat(lambda s: s.global_id, i32)
        check(lib().sofia_put_state(h, ptr(a), ptr(q), ptr(v), ptr(gid)))
        n1 = cat(lambda s: s.num1bond, i32)
        n2 = cat(lambda s: s.num2bond, i32)
        n3 = cat(lambda s: s.num3bond, i32)
        sp = cat(lambda s: s.specbond, i32)
        check(lib().sofia_set_exclusions(h, ptr(n1), ptr(n2), ptr(n3), ptr(sp)))
        per = lambda f: _c([len(f(r.system)) for r in reps], i32)
        bl = cat(lambda s: s.bondlist, i32)
        al = cat(lambda s: s.anglelist, i32)
        dl = cat(lambda s: s.dihedlist, i32)
        nb, na, nd = per(lambda s: s.bondlist), per(lambda s: s.anglelist), per(lambda s: s.dihedlist)
        check(lib().sofia_set_bonded(h, ptr(bl), len(bl), ptr(nb), ptr(al), len(al), ptr(na),
                                     ptr(dl), len(dl), ptr(nd)))
        if with_bonded:
            tl = [torsion_list(r.system) for r in reps]
            self._keep += _bonded_coeffs(h, np.concatenate(tl) if tl else np.zeros((0, 5), i32),
                                         [len(t) for t in tl])
            self._keep += _impropers(h, [r.system.improlist for r in reps], coul_flag)
        sn = cat(lambda s: s.shake_num, i32)
        sa = cat(lambda s: s.shake_atoms, i32)
        sb = cat(lambda s: s.shake_bond, f64)
        ns = per(lambda s: s.shake_num)
        check(lib().sofia_set_shake(h, len(sn), ptr(ns), ptr(sn), ptr(sa), ptr(sb)))
        mol = cat(lambda s: np.where(s.moltype == 1, s.mol, -1), i32)
        check(lib().sofia_set_molecules(h, ptr(mol)))
        th = (SofiaThermostat * len(reps))(*[SofiaThermostat(r.temp, r.gamma_ps, r.seed) for r in reps])
        check(lib().sofia_set_thermostat(h, th))
        if tuning:
            self.set_tuning(*tuning)

    def close(self):
        if getattr(self, "ctx", None):
            lib().sofia_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, lanes: int, block: int, list_waves: int):
        check(lib().sofia_set_tuning(self.ctx, lanes, block, list_waves))

    def tune_key(self) -> str:
        """The autotune cache bucket of this engine: precision, electrostatics, replica
        count and total atoms (both on a half-octave scale) and the mean number density
        (to 0.002 atoms/A^3)."""
        vol = float(sum(float(r.system.box) ** 3 for r in self.reps))
        rho = self.N / vol
        return (f"p{self.params.precision}c{self.params.coul_flag}"
                f"r{round(2 * np.log2(len(self.reps)))}n{round(2 * np.log2(self.N))}"
                f"d{round(rho / 0.002)}")

    def autotune(self, steps: int = 80, skins=(1.0, 1.3, 1.6, 2.0, 2.5), lanes=(0, 8, 16, 32),
                 waves=(4, 8, 16), graph_steps=(1, 20), cache: bool = True):
        """Pick the Verlet skin and the pair kernel's lanes per atom on the current state.
        First the list build's waves per workgroup on full rebuilds; then, for each skin, the
        lanes by timing the pair kernel alone, and the skins compared on whole MD steps (list
        length against rebuild rate); then the bonded evaluation (per-kind kernels or
        molecule bundles) on eager steps, and the MD steps per captured graph on whole
        steps.  The knobs change speed only.  Results are cached per bucket (tune_key: size
        and density on a coarse scale) for the process and, when SOFIA_TUNE_CACHE names a
        JSON file, across processes; a cached bucket is applied without timing.  Returns
        (skin, lanes, ms/step, list waves, bonded mode, steps per graph)."""
        key = self.tune_key()
        hit = _tune_cache_get(key) if cache else None
        if hit is not None:
            sk, L, ms, w, bmode, gk = hit
            self.set_skin(sk)
            self.set_tuning(L, PAIR_BLOCK, w)
            self.set_bonded_mode(bmode)
            self.set_graph_steps(gk)
            self.tune_log = {"cached": key}
            return sk, L, ms, w, bmode, gk
        import time
        best = None
        self.tune_log = {"pair_ms": {}, "step_ms": {}, "list_ms": {}}
        # the list build's waves first (full rebuilds at the middle skin), so that the skins
        # are compared with the rebuild cost they will run with
        self.set_skin(skins[len(skins) // 2])
        t_w = {}
        for w in waves:
            self.set_tuning(16, PAIR_BLOCK, w)
            t_w[w] = self.profile_step(3, full_rebuild=True)["list build"]
        self.tune_log["list_ms"] = t_w
        w = min(t_w, key=t_w.get)
        for sk in skins:
            self.set_skin(sk)
            t_l = {}
            for L in lanes:
                self.set_tuning(L, PAIR_BLOCK, w)
                t_l[L] = self.profile_nonbond(8)[0]
            self.tune_log["pair_ms"][sk] = t_l
            L = min(t_l, key=t_l.get)
            self.set_tuning(L, PAIR_BLOCK, w)
            self.run(5)
            self.synchronize()
            t0 = time.perf_counter()
            self.run(steps)
            self.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            self.tune_log["step_ms"][sk] = ms
            if best is None or ms < best[2]:
                best = (sk, L, ms)
        self.set_skin(best[0])
        self.set_tuning(best[1], PAIR_BLOCK, w)
        t_b = {}
        for mode in (0, 1):
            self.set_bonded_mode(mode)
            t_b[mode] = self.profile_step(6)["bonded"]
        self.tune_log["bonded_ms"] = t_b
        bmode = min(t_b, key=t_b.get)
        self.set_bonded_mode(bmode)
        t_g = {}
        for gk in graph_steps:
            self.set_graph_steps(gk)
            self.run(max(5, gk + 1))
            self.synchronize()
            t0 = time.perf_counter()
            self.run(steps)
            self.synchronize()
            t_g[gk] = 1e3 * (time.perf_counter() - t0) / steps
        self.tune_log["graph_ms"] = t_g
        gk = min(t_g, key=t_g.get)
        self.set_graph_steps(gk)
        out = (best[0], best[1], min(best[2], t_g[gk]), w, bmode, gk)
        if cache:
            _tune_cache_put(key, out)
        return out

    def set_graph_steps(self, k: int):
        """MD steps captured per hipGraph (launch overhead only)."""
        check(lib().sofia_set_graph_steps(self.ctx, int(k)))

    def set_list_headroom(self, headroom: float):
        """List stride = headroom x the largest neighbour count at list set-up."""
        check(lib().sofia_set_list_headroom(self.ctx, float(headroom)))

    def list_stats(self):
        """(list stride, segments redone after an in-segment list overflow)."""
        na, redos = C.c_int32(), C.c_int32()
        check(lib().sofia_md_list_stats(self.ctx, C.byref(na), C.byref(redos)))
        return na.value, redos.value

    def set_bonded_mode(self, mode: int):
        """Bonded items by kind (0, one thread per item) or by molecule bundles (1); changes
        speed only."""
        check(lib().sofia_set_bonded_mode(self.ctx, int(mode)))

    def set_skin(self, skin: float):
        """Engine Verlet skin (A); changes speed only, never the trajectory."""
        check(lib().sofia_set_skin(self.ctx, float(skin)))

    def set_thermostat(self, temps=None, gammas=None, seeds=None):
        """Per-replica Langevin temperature (K), friction (1/ps) and seed; None keeps the
        replica's spec.  A replica's noise stream stays keyed by (seed, step, atom)."""
        n = len(self.reps)
        temps = [r.temp for r in self.reps] if temps is None else list(temps)
        gammas = [r.gamma_ps for r in self.reps] if gammas is None else list(gammas)
        seeds = [r.seed for r in self.reps] if seeds is None else list(seeds)
        th = (SofiaThermostat * n)(*[SofiaThermostat(float(t), float(g), int(s))
                                     for t, g, s in zip(temps, gammas, seeds)])
        check(lib().sofia_set_thermostat(self.ctx, th))

    def temperatures(self) -> np.ndarray:
        """Kinetic temperature per replica (K): 2 KE / (k_B (3N - constraints - 3))."""
        ke = self.energies()[:, 7]
        dof = np.array([3 * r.system.np - len(S.constraint_pairs(r.system)[0]) - 3
                        for r in self.reps], float)
        return 2.0 * ke / (float(S.KBOLTZ) * dof)

    def init(self):
        check(lib().sofia_md_init(self.ctx))

    def minimize(self, nsteps: int = 200, gamma: float = 0.002, max_disp: float = 0.05):
        check(lib().sofia_minimize(self.ctx, nsteps, gamma, max_disp))

    def run(self, nsteps: int, use_graph: bool = True):
        check(lib().sofia_md_run(self.ctx, nsteps, 1 if use_graph else 0))

    def launch(self, nsteps: int, use_graph: bool = True):
        """Enqueue nsteps steps and return at once (finish() waits); engines launched back to
        back run concurrently on their own streams."""
        check(lib().sofia_md_launch(self.ctx, nsteps, 1 if use_graph else 0))

    def finish(self):
        check(lib().sofia_md_finish(self.ctx))

    def synchronize(self):
        check(lib().sofia_synchronize(self.ctx))

    def energies(self) -> np.ndarray:
        """(nrep, 9): potential, e_coul, pot_14, e_coul_14, e_bond, e_angle, e_dihedral, ke,
        e_improper."""
        out = np.empty(len(self.reps) * NENERGY, f64)
        check(lib().sofia_md_energies(self.ctx, ptr(out)))
        return out.reshape(-1, NENERGY)

    def rebuilds(self) -> np.ndarray:
        out = np.empty(len(self.reps), i32)
        check(lib().sofia_md_rebuilds(self.ctx, ptr(out)))
        return out

    def _by_gid(self, arrays):
        """Split device-order (N, 3) arrays per replica, each re-ordered by global id."""
        a = np.empty((self.N, 4), f32)
        gid = np.empty(self.N, i32)
        check(lib().sofia_get_state(self.ctx, ptr(a), None, ptr(gid)))
        out = []
        off = 0
        for n in self.natoms:
            sl = slice(off, off + n)
            order = np.argsort(gid[sl])
            out.append(tuple(x[sl][order] for x in arrays))
            off += n
        return out

    def state(self):
        """FP64 positions/velocities per replica, in each replica's own original atom
        order (global_id order)."""
        x = np.empty((self.N, 3), f64)
        v = np.empty((self.N, 3), f64)
        check(lib().sofia_get_state_f64(self.ctx, ptr(x), ptr(v)))
        return self._by_gid((x, v))

    def write_restart(self, k: int, path, num: int = 0, timeclock: float = 0.0):
        """Replica k's positions and velocities as a restart file in the layout of
        dump_global_velocity_id_mpi (Source/mod_datadump.f90:113-151), in global-id order."""
        from . import io as sio
        x, v = self.state()[k]
        s = self.reps[k].system
        sio.write_restart(path, float(s.box), float(self.reps[k].temp), num, timeclock,
                          np.arange(1, s.np + 1, dtype=i32), x, v)

    def forces(self):
        """Total force per replica after the last step (global_id order): the force the
        next step's first half-kick uses."""
        f = np.empty((self.N, 3), f64)
        check(lib().sofia_get_forces_f64(self.ctx, ptr(f)))
        return [t[0] for t in self._by_gid((f,))]

    def pair_forces(self):
        """The pair kernel alone at the current state: forces (N, 3) in device order and,
        per replica, (LJ energy, Coulomb energy, virial x, y, z)."""
        f = np.empty((self.N, 3), f64)
        e = np.empty((len(self.reps), 5), f64)
        check(lib().sofia_md_pair_forces(self.ctx, ptr(f), ptr(e)))
        return f, e

    def device_view(self):
        """FP64 positions (N, 3) and global ids (N,) in the current device order."""
        x = np.empty((self.N, 3), f64)
        gid = np.empty(self.N, i32)
        check(lib().sofia_get_state_f64(self.ctx, ptr(x), None))
        check(lib().sofia_get_state(self.ctx, None, None, ptr(gid)))
        return x, gid

    def replica_list(self, r: int):
        """Replica r's Verlet list: (nlist flat, numneigh, stride), 1-based ids within the
        replica in device order."""
        na = C.c_int32()
        n = int(self.natoms[r])
        check(lib().sofia_get_nlist_replica(self.ctx, r, None, None, C.byref(na)))
        nl = np.empty(n * na.value, i32)
        nn = np.empty(n, i32)
        check(lib().sofia_get_nlist_replica(self.ctx, r, ptr(nl), ptr(nn), C.byref(na)))
        return nl, nn, na.value

    def config(self) -> dict:
        lanes, waves, bmode, rec = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        skin = C.c_double()
        check(lib().sofia_get_engine_config(self.ctx, C.byref(lanes), C.byref(waves), C.byref(bmode),
                                            C.byref(rec), C.byref(skin)))
        return {"lanes": lanes.value, "list_waves": waves.value, "bonded_mode": bmode.value,
                "record_bytes": rec.value, "skin": skin.value}

    def put_state(self, states):
        """Upload FP64 (x, v) per replica, each in its own global_id order; valid before
        the first sort (a fresh engine holds every replica in that order)."""
        x = np.ascontiguousarray(np.concatenate([s[0] for s in states]), dtype=f64)
        v = np.ascontiguousarray(np.concatenate([s[1] for s in states]), dtype=f64)
        check(lib().sofia_put_state_f64(self.ctx, ptr(x), ptr(v)))

    def steps(self) -> np.ndarray:
        out = np.empty(len(self.reps), np.uint64)
        check(lib().sofia_get_replica_steps(self.ctx, ptr(out)))
        return out

    def set_steps(self, steps):
        s = np.ascontiguousarray(steps, dtype=np.uint64)
        check(lib().sofia_set_replica_steps(self.ctx, ptr(s)))

    def resume(self, forces=None):
        """Continue from put_state/set_steps.  `forces` (per replica, global_id order, or
        None for a replica whose forces are to be recomputed) carries each replica's last
        total force over, which makes the continuation bit-exact."""
        if forces is None or all(f is None for f in forces):
            check(lib().sofia_md_resume(self.ctx))
            return
        have = _c([f is not None for f in forces], i32)
        f = np.ascontiguousarray(np.concatenate(
            [f if f is not None else np.zeros((n, 3)) for f, n in zip(forces, self.natoms)]), dtype=f64)
        check(lib().sofia_md_resume_forces(self.ctx, ptr(f), ptr(have)))

    def profile_nonbond(self, launches: int = 20):
        """(ms per pair-kernel launch, list entries inside the cutoff)."""
        ms = C.c_double()
        pairs = C.c_int64()
        check(lib().sofia_profile_nonbond(self.ctx, launches, C.byref(ms), C.byref(pairs)))
        return ms.value, pairs.value

    def profile_copy(self, nbytes: int = 1 << 30, launches: int = 10) -> float:
        """ms per launch of a plain device copy of nbytes (read nbytes + write nbytes)."""
        ms = C.c_double()
        check(lib().sofia_profile_copy(self.ctx, int(nbytes), launches, C.byref(ms)))
        return ms.value

    PHASES = ("half-kick", "drift", "check", "sort+gather", "list build", "rebuild flags",
              "zero forces", "pair kernel", "bonded", "langevin", "shake", "half-kick 2")

    def profile_step(self, steps: int = 50, full_rebuild: bool = False) -> dict:
        """Average ms per step of each phase of an MD step (eager launches, HIP events);
        full_rebuild makes every replica rebuild its list on every profiled step."""
        out = np.zeros(len(self.PHASES), f64)
        check(lib().sofia_profile_step(self.ctx, steps, 1 if full_rebuild else 0, ptr(out)))
        return dict(zip(self.PHASES, out.tolist()))

    def greatest_cluster(self, denscut: float, densmin: int):
        """get_greatestCluster per replica over the solute molecules: (total, core,
        surface) arrays (Source/mod_cluster.f90:15-83)."""
        n = len(self.reps)
        tot, core, surf = np.empty(n, i32), np.empty(n, i32), np.empty(n, i32)
        check(lib().sofia_cluster_greatest(self.ctx, float(denscut), int(densmin), ptr(tot), ptr(core),
                                           ptr(surf)))
        return tot, core, surf

    def largest_cluster(self, rcontact: float = 3.5) -> np.ndarray:
        o = _c([4, 5], i32)
        nt = _c([3], i32)
        out = np.empty(len(self.reps), i32)
        check(lib().sofia_cluster_largest(self.ctx, float(rcontact), len(o), ptr(o), len(nt), ptr(nt),
                                          ptr(out)))
        return out
