"""Glycine/water systems in Sofia's data conventions, plus the derived-constant contract.

Everything here is host-side numpy: the force field table, the per-molecule topology
templates, the exclusion table (``specbond``), the bonded and SHAKE lists, and a seeded
synthetic-system generator (jittered lattice, uniform SO(3) orientations, Maxwell-Boltzmann
velocities).  Arrays follow the reference layout:

* atom ids stored in lists are 1-based (``Source/global.f90:99-110, 430-440``);
* ``specbond`` is (np, 12) here, i.e. the transpose of Fortran's column-major
  ``specbond(12, np)``, so row ``i`` is column ``i`` of the reference array;
* ``num1bond <= num2bond <= num3bond`` are cumulative counts
  (``Source/mod_buildbond_lists.f90:99-101``);
* positions, velocities and charges are float32 (``real*4``), the pair cutoffs and
  DSF constants are float64 after the same float32 rounding the reference applies
  (``Source/get_started.f90:25-117``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32
f64 = np.float64

# --------------------------------------------------------------------------------------
# Constants (Source/get_started.f90:32-38, 102-110).  The reference assigns single
# precision literals, so each value is the float32 rounding of the literal.
# --------------------------------------------------------------------------------------
KCAL2AMU = f32(0.0004184)            # real*4 (get_started.f90:32)
AMU2E = f64(f32(2390.057361))        # double, from a real*4 literal (:33)
COULPRE = f32(332.06371)             # real*4 (:34)
NKTV2P = f64(f32(68568.415))         # (:35)
KBOLTZ = f32(0.001987204181)         # real*4 from a d0 literal (:38)
EWALD_F = f64(f32(1.12837917))       # (:102)
EWALD_P = f64(f32(0.3275911))        # (:103)
A1 = f64(f32(0.254829592))           # (:104-108), Abramowitz & Stegun 7.1.26
A2 = f64(f32(-0.284496736))
A3 = f64(f32(1.421413741))
A4 = f64(f32(-1.453152027))
A5 = f64(f32(1.061405429))
MY_PIS = f64(f32(1.77245385090551602729))   # no d0 suffix (:109)
MY_PIS_INV = f64(1.0) / MY_PIS               # (:110)
SHAKE_TOL = f64(f32(0.0001))                 # tolerance (:71)
SHAKE_MAXITER = 25                           # maxiter (:70)
NEIGH_ALLOC0 = 512                           # Source/mod_memory.f90:137
SPEC_SLOTS = 12                              # specbond(12, np), Source/mod_memory.f90:33

# --------------------------------------------------------------------------------------
# Force field: input_files/topol_cndo.top:1-87 (atom types, bonds, angles, torsions,
# improper, SHAKE clusters).  Columns: name, rmin/2 (A), eps (kcal/mol), q (e), mass (amu).
# --------------------------------------------------------------------------------------
ATOM_TYPES = [
    ("C1", 1.9080, 0.0860, 0.374, 12.011),
    ("C2", 1.9080, 0.1094, -0.021, 12.011),
    ("N", 1.8240, 0.1700, 0.022, 14.007),
    ("O1", 1.6612, 0.2100, -0.526, 15.999),
    ("O2", 1.6612, 0.2100, -0.483, 15.999),
    ("H1", 0.6000, 0.0157, 0.164, 1.008),
    ("H2", 0.6000, 0.0157, 0.208, 1.008),
    ("H3", 0.6000, 0.0157, 0.199, 1.008),
    ("H4", 1.1000, 0.0157, 0.033, 1.008),
    ("H5", 1.1000, 0.0157, 0.030, 1.008),
    ("OW", 1.77685472, 0.1554, -0.8476, 15.99900),
    ("HW", 0.0000, 0.0000, 0.4238, 1.00800),
]
NTYPES = len(ATOM_TYPES)

# bond types: (kb, r0);  angle types: (k, theta0 deg);  torsion types: (K, n, shift deg)
BOND_COEFF = [(328.30, 1.5080), (648.00, 1.2140), (293.60, 1.4990), (338.70, 1.0910),
              (369.00, 1.0330), (600.00, 1.0000)]
ANGLE_COEFF = [(68.03, 123.11), (78.17, 130.38), (65.07, 114.21), (47.23, 109.54),
               (49.02, 107.91), (39.04, 110.74), (46.19, 110.11), (40.52, 108.11),
               (75.00, 109.47)]
DIHED_COEFF = [(0.0000, 2, 180), (0.1556, 3, 0)]
IMPRO_COEFF = [(1.1000, -1, 2)]
# improper types as the engine takes them: (K kcal/mol/rad^2, chi0 rad).  The reference
# reads K from the line above but sets improcoeff(2, :) from `theta0`, a local that the
# read never assigns (Source/mod_readfile.f90:230, 427-433), so the file fixes no chi0.
# With the atom order (C-alpha, O1, C1 central, O2) a planar carboxylate has chi = pi
# (C-alpha and O2 lie on opposite sides of the O1-C1 axis), so chi0 = pi makes the term
# a planarity restraint; it also keeps the cusp of (acos(c) - chi0)^2 at chi = 0, far
# from the minimum, where chi0 = 0 would put it at the planar geometry.  The reference's
# minimised input input_files/glycine_1500_solv_min.gro has chi = 166-180 deg (mean 175).
IMPRO_K_CHI0 = [(1.1000, math.pi)]


@dataclass
class MolTemplate:
    """One molecule type: atom types (1-based), local bonded lists (1-based), SHAKE."""
    name: str
    types: list
    bonds: list      # (a1, a2, btype) in topology order
    angles: list     # (a1, a2(middle), a3, atype)
    diheds: list     # (a1, a2, a3, a4, dtype)
    impros: list     # (a1, a2, a3(central), a4, itype)
    shake: list      # (num, atoms tuple, bond lengths tuple)
    geom: np.ndarray  # (natom, 3) reference geometry, A


def _glycine_geometry() -> np.ndarray:
    # Heavy-atom frame from input_files/mol_cndo_alpha.input:5-14; each H is moved along
    # its bond to the SHAKE length of topol_cndo.top:83-85 so constraints hold exactly.
    g = np.array([[0.250, 1.493, 0.310], [0.766, 1.735, -1.000], [2.064, 1.075, -1.216],
                  [1.080, 1.128, 1.107], [-0.972, 1.694, 0.503], [2.334, 1.197, -2.020],
                  [2.663, 1.425, -0.766], [2.003, 0.239, -1.113], [0.874, 2.633, -1.076],
                  [0.149, 1.413, -1.546]], dtype=f64)
    for h, parent, d in ((5, 2, 1.0330), (6, 2, 1.0330), (7, 2, 1.0330), (8, 1, 1.0910),
                         (9, 1, 1.0910)):
        u = g[h] - g[parent]
        g[h] = g[parent] + u * (d / np.linalg.norm(u))
    return g - g.mean(axis=0)


def _water_geometry() -> np.ndarray:
    # O-H 1.0 A, H-H 1.633 A (topol_cndo.top:87)
    half = 0.5 * 1.633
    h = math.sqrt(1.0 - half * half)
    g = np.array([[0.0, 0.0, 0.0], [half, h, 0.0], [-half, h, 0.0]], dtype=f64)
    return g - g.mean(axis=0)


GLYCINE = MolTemplate(
    name="glycine",
    types=list(range(1, 11)),
    bonds=[(1, 2, 1), (1, 4, 2), (1, 5, 2), (2, 3, 3), (2, 9, 4), (2, 10, 4), (3, 6, 5),
           (3, 7, 5), (3, 8, 5)],
    angles=[(2, 1, 4, 1), (2, 1, 5, 1), (4, 1, 5, 2), (1, 2, 3, 3), (1, 2, 9, 4),
            (1, 2, 10, 4), (3, 2, 9, 5), (3, 2, 10, 5), (9, 2, 10, 6), (2, 3, 6, 7),
            (2, 3, 7, 7), (2, 3, 8, 7), (6, 3, 7, 8), (6, 3, 8, 8), (7, 3, 8, 8)],
    diheds=[(4, 1, 2, 3, 1), (5, 1, 2, 3, 1), (4, 1, 2, 9, 1), (4, 1, 2, 10, 1),
            (5, 1, 2, 9, 1), (5, 1, 2, 10, 1), (1, 2, 3, 6, 2), (1, 2, 3, 7, 2),
            (1, 2, 3, 8, 2), (9, 2, 3, 6, 2), (9, 2, 3, 7, 2), (9, 2, 3, 8, 2),
            (10, 2, 3, 6, 2), (10, 2, 3, 7, 2), (10, 2, 3, 8, 2)],
    impros=[(2, 4, 1, 5, 1)],
    shake=[(3, (2, 9, 10), (1.0910, 1.0910, 0.0)),
           (4, (3, 6, 7, 8), (1.0330, 1.0330, 1.0330))],
    geom=_glycine_geometry(),
)

WATER = MolTemplate(
    name="water",
    types=[11, 12, 12],
    bonds=[(1, 3, 6), (1, 2, 6)],
    angles=[(2, 1, 3, 9)],
    diheds=[],
    impros=[],
    shake=[(-3, (1, 2, 3), (1.0000, 1.0000, 1.6330))],
    geom=_water_geometry(),
)


def mol_specbond(t: MolTemplate):
    """Per-atom exclusion slots of one molecule, filled the way read_topol does.

    Slots 1..n1 hold 1-2 partners in bond order, then 1-3 partners in angle order, then
    1-4 partners in torsion order (Source/mod_readfile.f90:301-392); counts are made
    cumulative as in Source/mod_buildbond_lists.f90:99-101.
    """
    n = len(t.types)
    n1 = [0] * n
    n2 = [0] * n
    n3 = [0] * n
    b12 = [[] for _ in range(n)]
    b13 = [[] for _ in range(n)]
    b14 = [[] for _ in range(n)]
    for a1, a2, _ in t.bonds:
        b12[a1 - 1].append(a2)
        b12[a2 - 1].append(a1)
    for a1, _, a3, _ in t.angles:
        b13[a1 - 1].append(a3)
        b13[a3 - 1].append(a1)
    for a1, _, _, a4, _ in t.diheds:
        b14[a1 - 1].append(a4)
        b14[a4 - 1].append(a1)
    spec = np.zeros((n, SPEC_SLOTS), dtype=np.int32)
    for i in range(n):
        row = b12[i] + b13[i] + b14[i]
        assert len(row) <= SPEC_SLOTS
        spec[i, :len(row)] = row
        n1[i] = len(b12[i])
        n2[i] = n1[i] + len(b13[i])
        n3[i] = n2[i] + len(b14[i])
    return np.array(n1, np.int32), np.array(n2, np.int32), np.array(n3, np.int32), spec


# --------------------------------------------------------------------------------------
# Parameters and derived constants
# --------------------------------------------------------------------------------------
@dataclass
class Params:
    """Run parameters with the reference's storage precision.

    ``rcut``, ``cut_coul``, ``pad``, ``dt``, ``temp`` are real*4; ``alpha``, ``gewald``
    and every derived cutoff square are real*8 computed as Source/get_started.f90:45-117.
    """
    rcut: float = 10.0
    cut_coul: float = 10.0
    pad: float = 2.0
    alpha: float = 0.2
    gewald: float = 0.0
    dt: float = 2.0
    temp: float = 300.0
    gamma_ps: float = 1.0          # Langevin friction, 1/ps
    fudge_12: float = 0.0
    fudge_13: float = 0.0
    fudge_14_coul: float = 1.0 / 1.2   # Source/mod_readfile.f90:100

    def derived(self, box: float) -> dict:
        rcut = f32(self.rcut)
        cut_coul = f32(self.cut_coul)
        pad = f32(self.pad)
        box32 = f32(box)
        alpha = f64(self.alpha)
        cut_coulsq = f64(f32(cut_coul * cut_coul))       # real*4 product, :75
        rcut2 = f64(f32(rcut * rcut))                    # :74
        rv = f64(f32(pad + cut_coul))                    # real*4 sum, :45
        rv2 = rv * rv                                     # :46
        rvrc2 = f64(pad) * f64(pad)                       # :47-48
        ibox = f32(f64(1.0) / f64(box32))                 # :66
        cc = f64(cut_coul)
        erfcc = math.erfc(alpha * cc)                     # :112
        erfcd = math.exp(-alpha * alpha * cc * cc)        # :113
        f_shift = -(erfcc / cut_coulsq + ((2.0 / MY_PIS) * alpha) * erfcd / cc)   # :115
        e_shift = erfcc / cc - f_shift * cc               # :116
        dt = f32(self.dt)
        dthalf = f32(f64(0.5) * f64(dt) * f64(KCAL2AMU))  # :52
        return dict(rcut=rcut, cut_coul=cut_coul, pad=pad, box=box32, ibox=ibox,
                    alpha=alpha, gewald=f64(self.gewald), rcut2=rcut2, cut_coulsq=cut_coulsq,
                    rv=rv, rv2=rv2, rvrc2=rvrc2, f_shift=f64(f_shift), e_shift=f64(e_shift),
                    dt=dt, dthalf=dthalf)


def lj_tables():
    """lj1..lj4 (real*4) and lj14_1..4 (real*8), Source/mod_readfile.f90:595-615.

    Lorentz-Berthelot on Rmin/2: Rmin = sig_i + sig_j, eps = sqrt(eps_i eps_j).
    Returned flat with index (itype-1)*ntypes + (jtype-1).
    """
    sig = np.array([t[1] for t in ATOM_TYPES], dtype=f64)
    eps = np.array([t[2] for t in ATOM_TYPES], dtype=f64)
    le = np.sqrt(np.outer(eps, eps))
    ls = sig[:, None] + sig[None, :]
    s6 = ls ** 6
    s12 = ls ** 12
    lj1 = (12.0 * le * s12).astype(f32).ravel()
    lj2 = (12.0 * le * s6).astype(f32).ravel()
    lj3 = (le * s12).astype(f32).ravel()
    lj4 = (2.0 * le * s6).astype(f32).ravel()
    lj14_1 = (0.5 * 12.0 * le * s12).ravel()
    lj14_2 = (0.5 * 12.0 * le * s6).ravel()
    lj14_3 = (0.5 * le * s12).ravel()
    lj14_4 = (le * s6).ravel()
    return lj1, lj2, lj3, lj4, lj14_1, lj14_2, lj14_3, lj14_4


def type_mass() -> np.ndarray:
    return np.array([t[4] for t in ATOM_TYPES], dtype=f64)


def type_charge() -> np.ndarray:
    return np.array([t[3] for t in ATOM_TYPES], dtype=f64)


def cell_geometry(box: float, rcut: float, cut_coul: float):
    """check_allocation's cell grid (Source/mod_neighbor.f90:1446-1451), in real*4.

    rn = box/int(box/max(rcut, cut_coul)); ncellT = nint(vol/rn**3);
    ncellD = nint(ncellT**(1/3)) with vol = box**3 (Source/get_started.f90:64).
    """
    box = f32(box)
    rc = f32(cut_coul) if f32(cut_coul) > f32(rcut) else f32(rcut)
    n = int(f32(box / rc))
    rn = f32(box / f32(n))
    vol = f32(f32(box * box) * box)
    rn3 = f32(f32(rn * rn) * rn)
    ncellT = int(_nint(float(f32(vol / rn3))))
    ncellD = int(_nint(float(ncellT) ** (1.0 / 3.0)))
    return rn, ncellT, ncellD


def _nint(x: float) -> int:
    """Fortran nint: round half away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


# --------------------------------------------------------------------------------------
# System container
# --------------------------------------------------------------------------------------
@dataclass
class SofiaSystem:
    box: np.float32
    pos: np.ndarray          # (np, 3) float32
    types: np.ndarray        # (np,) int32, 1-based
    q: np.ndarray            # (np,) float32
    v: np.ndarray            # (np, 3) float32
    global_id: np.ndarray    # (np,) int32, 1-based
    mol: np.ndarray          # (np,) int32 molecule index (0-based)
    moltype: np.ndarray      # (np,) int32: 1 glycine, 2 water
    num1bond: np.ndarray
    num2bond: np.ndarray
    num3bond: np.ndarray
    specbond: np.ndarray     # (np, 12) int32, 1-based
    bondlist: np.ndarray     # (nb, 3) int32
    anglelist: np.ndarray    # (na, 4)
    dihedlist: np.ndarray    # (nd, 5)
    improlist: np.ndarray    # (ni, 5)
    shake_num: np.ndarray    # (ns,) int32  (1, 3, 4, -3)
    shake_atoms: np.ndarray  # (ns, 4) int32 1-based (0 = unused)
    shake_bond: np.ndarray   # (ns, 3) float64
    n_gly: int = 0
    n_wat: int = 0
    seed: int = 0
    params: Params = field(default_factory=Params)

    @property
    def np(self) -> int:
        return int(self.pos.shape[0])

    def ainfo(self) -> np.ndarray:
        """(np, 4) int32-viewable AoS matching ``type ainfo`` (16 B/atom)."""
        a = np.empty((self.np, 4), dtype=np.float32)
        a[:, :3] = self.pos
        a[:, 3] = self.types.astype(np.int32).view(np.float32)
        return a

    def derived(self) -> dict:
        return self.params.derived(float(self.box))

    def copy(self) -> "SofiaSystem":
        import copy as _c
        return _c.deepcopy(self)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    # Uniform SO(3): normalised quaternion from four Gaussians.
    qv = rng.standard_normal((n, 4))
    qv /= np.linalg.norm(qv, axis=1, keepdims=True)
    w, x, y, z = qv.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def box_for(n_gly: int, n_wat: int, density_g_cm3: float = 1.0) -> float:
    mass = type_mass()
    m_gly = mass[:10].sum()
    m_wat = mass[10] + 2 * mass[11]
    total = n_gly * m_gly + n_wat * m_wat            # amu
    vol = total * 1.66053906660 / density_g_cm3       # A^3
    return vol ** (1.0 / 3.0)


def make_system(n_gly: int, n_wat: int, seed: int = 20261202, params: Params | None = None,
                box: float | None = None, temp: float | None = None,
                jitter: float = 0.2) -> SofiaSystem:
    """Seeded synthetic glycine/water box at 1.0 g/cm^3.

    Glycine centres take a sparse sub-lattice of a cubic lattice, waters the lattice sites
    that are at least 2.9 A from every glycine atom; centres are jittered, orientations
    are uniform on SO(3) and re-drawn for molecules in a clash (heavy atoms closer than
    2.4 A or any atoms closer than 1.5 A).  Velocities are Maxwell-Boltzmann at ``temp``
    with the centre-of-mass drift removed and the constraint components projected out.
    """
    from scipy.spatial import cKDTree
    params = params or Params()
    if temp is not None:
        params.temp = temp
    rng = np.random.default_rng(seed)
    nmol = n_gly + n_wat
    L = float(box) if box is not None else box_for(n_gly, n_wat)
    heavy_types = np.array([1, 2, 3, 4, 5, 11])

    def lattice(nl):
        a = L / nl
        idx = np.arange(nl ** 3)
        grid = np.stack([idx % nl, (idx // nl) % nl, idx // (nl * nl)], axis=1)
        return grid, (grid + 0.5) * a, a

    # lattice with one site per water and ~5 per glycine (its volume is ~4 waters)
    nl = max(3, int(math.ceil((n_wat + 5 * n_gly) ** (1.0 / 3.0))))
    grid, centres, a = lattice(nl)
    stride = 3 if (nl // 3) ** 3 >= n_gly else 2
    sub = np.where(np.all(grid % stride == 0, axis=1))[0]
    if n_gly > len(sub):
        raise ValueError("too many glycines for the lattice")
    gsites = rng.choice(sub, size=n_gly, replace=False) if n_gly else np.zeros(0, int)
    free = np.ones(len(centres), bool)
    free[gsites] = False
    if n_gly:
        # the 4 nearest lattice neighbours of each glycine site are its own volume
        tree = cKDTree(centres, boxsize=L)
        _, nb = tree.query(centres[gsites], k=5)
        free[nb.ravel()] = False
    fidx = np.where(free)[0]
    if n_wat > len(fidx):
        raise ValueError("lattice too small for the waters")
    wsites = rng.choice(fidx, size=n_wat, replace=False)
    C = np.concatenate([centres[gsites], centres[wsites]]) + rng.uniform(-jitter, jitter, (nmol, 3))
    R = _random_rotations(rng, nmol)

    natom_mol = np.array([10] * n_gly + [3] * n_wat)
    offsets = np.concatenate([[0], np.cumsum(natom_mol)])
    N = int(offsets[-1])
    mol_of_atom = np.repeat(np.arange(nmol), natom_mol)
    types = np.concatenate([np.array(GLYCINE.types * n_gly, np.int32),
                            np.array(WATER.types * n_wat, np.int32)]) if nmol else np.zeros(0, np.int32)
    is_heavy = np.isin(types, heavy_types)
    local = np.concatenate([np.tile(GLYCINE.geom, (n_gly, 1)), np.tile(WATER.geom, (n_wat, 1))])

    def place():
        return C[mol_of_atom] + np.einsum("nij,nj->ni", R[mol_of_atom], local)

    # rigid-body overlap removal: molecules in contact are pushed apart along the contact
    # (heavy-heavy below 2.5 A, any pair below 1.7 A); clashing waters are also re-rotated
    xyz = place()
    for it in range(150):
        tree = cKDTree(np.mod(xyz, L), boxsize=L)
        pr = tree.query_pairs(2.5, output_type="ndarray")
        if len(pr):
            pr = pr[mol_of_atom[pr[:, 0]] != mol_of_atom[pr[:, 1]]]
        if len(pr) == 0:
            break
        dv = xyz[pr[:, 0]] - xyz[pr[:, 1]]
        dv -= L * np.round(dv / L)
        dd = np.sqrt((dv * dv).sum(1))
        thr = np.where(is_heavy[pr[:, 0]] & is_heavy[pr[:, 1]], 2.5, 1.7)
        hit = dd < thr
        if not hit.any():
            break
        pr, dv, dd, thr = pr[hit], dv[hit], dd[hit], thr[hit]
        u = dv / np.maximum(dd, 1e-3)[:, None]
        push = (0.5 * (thr - dd) + 0.02)[:, None] * u
        D = np.zeros((nmol, 3))
        np.add.at(D, mol_of_atom[pr[:, 0]], push)
        np.add.at(D, mol_of_atom[pr[:, 1]], -push)
        norm = np.linalg.norm(D, axis=1, keepdims=True)
        D *= np.minimum(1.0, 0.3 / np.maximum(norm, 1e-12))
        C += D
        if it % 4 == 3:
            bad = np.unique(mol_of_atom[pr.ravel()])
            bad = bad[bad >= n_gly]
            R[bad] = _random_rotations(rng, len(bad))
        xyz = place()

    # wrap into (0, box] as pbc() does (Source/mod_pbc.f90:6-39)
    box32 = f32(L)
    pos = np.mod(xyz, L).astype(f32)
    pos = np.where(pos <= 0, pos + box32, pos).astype(f32)
    pos = np.where(pos > box32, pos - box32, pos).astype(f32)
    return assemble(pos, box32, n_gly, n_wat, seed, params, rng)


def assemble(pos: np.ndarray, box32, n_gly: int, n_wat: int, seed: int, params: Params,
             rng: np.random.Generator) -> SofiaSystem:
    """Topology, charges, SHAKE clusters and Maxwell-Boltzmann velocities for positions
    (float32, n_gly glycines of 10 atoms in template order, then n_wat waters O, H, H)."""
    nmol = n_gly + n_wat
    natom_mol = np.array([10] * n_gly + [3] * n_wat)
    offsets = np.concatenate([[0], np.cumsum(natom_mol)])
    N = int(offsets[-1])
    mol_of_atom = np.repeat(np.arange(nmol), natom_mol)
    types = np.concatenate([np.array(GLYCINE.types * n_gly, np.int32),
                            np.array(WATER.types * n_wat, np.int32)]) if nmol else np.zeros(0, np.int32)
    if len(pos) != N:
        raise ValueError(f"{len(pos)} positions for {n_gly} glycines and {n_wat} waters ({N} atoms)")
    templates = [GLYCINE] * n_gly + [WATER] * n_wat

    # ---- topology -------------------------------------------------------------------
    qt = type_charge()
    q = qt[types - 1].astype(f32)
    moltype = np.repeat(np.array([1] * n_gly + [2] * n_wat, np.int32), natom_mol)

    def tile(template: MolTemplate, count: int, base: int):
        n1, n2, n3, spec = mol_specbond(template)
        na = len(template.types)
        off = (base + np.arange(count) * na)[:, None]
        spec_g = np.where(spec[None, :, :] > 0, spec[None, :, :] + off[:, :, None], 0)
        return (np.tile(n1, count), np.tile(n2, count), np.tile(n3, count),
                spec_g.reshape(count * na, SPEC_SLOTS).astype(np.int32))

    parts = []
    if n_gly:
        parts.append(tile(GLYCINE, n_gly, 0))
    if n_wat:
        parts.append(tile(WATER, n_wat, 10 * n_gly))
    num1 = np.concatenate([p[0] for p in parts]).astype(np.int32)
    num2 = np.concatenate([p[1] for p in parts]).astype(np.int32)
    num3 = np.concatenate([p[2] for p in parts]).astype(np.int32)
    spec = np.concatenate([p[3] for p in parts]).astype(np.int32)

    bl, al, dl, il = _bonded_lists(spec, num1, num2, num3, n_gly, n_wat)

    # SHAKE clusters in molecule order (Source/mod_buildbond_lists.f90:464-653)
    sh_num, sh_atoms, sh_bond = [], [], []
    for m, t in enumerate(templates):
        base = offsets[m]
        for num, atoms, bonds in t.shake:
            sh_num.append(num)
            a4 = list(atoms) + [0] * (4 - len(atoms))
            sh_atoms.append([base + a if a else 0 for a in a4])
            sh_bond.append(bonds)
    shake_num = np.array(sh_num, np.int32)
    shake_atoms = np.array(sh_atoms, np.int32).reshape(-1, 4)
    shake_bond = np.array(sh_bond, f64).reshape(-1, 3)

    system = SofiaSystem(box=box32, pos=pos, types=types, q=q,
                         v=np.zeros((N, 3), f32), global_id=np.arange(1, N + 1, dtype=np.int32),
                         mol=mol_of_atom.astype(np.int32), moltype=moltype,
                         num1bond=num1, num2bond=num2, num3bond=num3, specbond=spec,
                         bondlist=bl, anglelist=al, dihedlist=dl, improlist=il,
                         shake_num=shake_num, shake_atoms=shake_atoms, shake_bond=shake_bond,
                         n_gly=n_gly, n_wat=n_wat, seed=seed, params=params)
    system.v = maxwell_boltzmann(system, params.temp, rng)
    return system


def _bonded_lists(spec, num1, num2, num3, n_gly, n_wat):
    """bondlist/anglelist/dihedlist/improlist from the exclusion slots.

    Pairs are kept with partner > self, walking atoms in order, as
    Source/mod_buildbond_lists.f90:262-461 does.
    """
    def from_templates(getter, width):
        out = []
        for t, count, base in ((GLYCINE, n_gly, 0), (WATER, n_wat, 10 * n_gly)):
            items = getter(t)
            if not items or count == 0:
                continue
            arr = np.array(items, np.int32)
            na = len(t.types)
            offs = base + np.arange(count, dtype=np.int32) * na
            rep = np.repeat(arr[None, :, :], count, axis=0).copy()
            rep[:, :, :width] += offs[:, None, None]
            out.append(rep.reshape(-1, arr.shape[1]))
        return np.concatenate(out) if out else np.zeros((0, width + 1), np.int32)

    # the lists come out in the reference's (atom, slot) walk order
    def order12(t):
        n1, n2, n3, sp = mol_specbond(t)
        btype = {}
        for a1, a2, bt in t.bonds:
            btype[(a1, a2)] = btype[(a2, a1)] = bt
        res = []
        for i in range(len(t.types)):
            for l in range(n1[i]):
                j = sp[i, l]
                if j > i + 1:
                    res.append((i + 1, j, btype[(i + 1, j)]))
        return res

    def order13(t):
        n1, n2, n3, sp = mol_specbond(t)
        mid = {}
        for a1, a2, a3, at in t.angles:
            mid[(a1, a3)] = mid[(a3, a1)] = (a2, at)
        res = []
        for i in range(len(t.types)):
            for l in range(n1[i], n2[i]):
                j = sp[i, l]
                if j > i + 1:
                    m, at = mid[(i + 1, j)]
                    res.append((i + 1, m, j, at))
        return res

    def order14(t):
        # read_topol stores the two middle atoms unreversed for both end atoms
        # (Source/mod_readfile.f90:390-398), so an entry seen from atom4 keeps (a2, a3).
        n1, n2, n3, sp = mol_specbond(t)
        mids = {}
        for a1, a2, a3, a4, dt in t.diheds:
            mids[(a1, a4)] = (a2, a3, dt)
            mids[(a4, a1)] = (a2, a3, dt)
        res = []
        for i in range(len(t.types)):
            for l in range(n2[i], n3[i]):
                j = sp[i, l]
                if j > i + 1:
                    m2, m3, dt = mids[(i + 1, j)]
                    res.append((i + 1, m2, m3, j, dt))
        return res

    bl = from_templates(order12, 2)
    al = from_templates(order13, 3)
    dl = from_templates(order14, 4)
    il = from_templates(lambda t: [tuple(x) for x in t.impros], 4)
    return bl, al, dl, il


def maxwell_boltzmann(system: SofiaSystem, temp: float, rng: np.random.Generator) -> np.ndarray:
    """Gaussian velocities (A/fs) at ``temp`` with KE in kcal/mol = 1/2 m v^2 amu2e
    (Source/mod_force.f90:727-740), COM drift removed, bond components projected out."""
    m = type_mass()[system.types - 1]
    sigma = np.sqrt(f64(KBOLTZ) * temp / (m * AMU2E))
    v = rng.standard_normal((system.np, 3)) * sigma[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    v = project_constraints(system, v)
    return v.astype(f32)


def constraint_pairs(system: SofiaSystem):
    """Flattened (a, b, d) constraint pairs (0-based) of every SHAKE cluster."""
    num = system.shake_num
    at = system.shake_atoms.astype(np.int64) - 1
    bd = system.shake_bond
    # (first atom slot, second atom slot, bond slot) per constraint of each cluster type
    table = {1: [(0, 1, 0)], 3: [(0, 1, 0), (0, 2, 1)], 4: [(0, 1, 0), (0, 2, 1), (0, 3, 2)],
             -3: [(0, 1, 0), (0, 2, 1), (1, 2, 2)]}
    a, b, d = [], [], []
    for kind, cons in table.items():
        sel = np.where(num == kind)[0]
        for p, q, k in cons:
            a.append(at[sel, p]); b.append(at[sel, q]); d.append(bd[sel, k])
    if not a:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, f64)
    return np.concatenate(a), np.concatenate(b), np.concatenate(d).astype(f64)


def project_constraints(system: SofiaSystem, v: np.ndarray, sweeps: int = 30) -> np.ndarray:
    """Remove relative velocity along each constrained bond (Jacobi-style RATTLE sweeps)."""
    a, b, _ = constraint_pairs(system)
    if len(a) == 0:
        return v
    m = type_mass()[system.types - 1]
    x = system.pos.astype(f64)
    L = float(system.box)
    r = x[a] - x[b]
    r -= L * np.round(r / L)
    r2 = (r * r).sum(1)
    ima, imb = 1.0 / m[a], 1.0 / m[b]
    v = v.astype(f64).copy()
    n = len(v)
    for _ in range(sweeps):
        dv = v[a] - v[b]
        g = (dv * r).sum(1) / (r2 * (ima + imb))
        corr = g[:, None] * r
        for c in range(3):
            v[:, c] += (np.bincount(a, -0.5 * corr[:, c] * ima, minlength=n) +
                        np.bincount(b, 0.5 * corr[:, c] * imb, minlength=n))
    return v


CONFIGS = {
    # name: (n_gly, n_wat)
    "small": (150, 4000),       # 13,500 atoms
    "plain": (1200, 64000),     # 204,000 atoms
    "mixed": (6000, 320000),    # 1,020,000 atoms
}


def replica_sizes(n: int = 64, lo: int = 5000, hi: int = 60000, seed: int = 20261205):
    """Batched-replica compositions: N log-uniform in [lo, hi], glycine mole fraction
    uniform in [0.01, 0.04]; returns a list of (n_gly, n_wat)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        x = rng.uniform(0.01, 0.04)
        # N = 10 g + 3 w, g/(g+w) = x  ->  g = x N / (10x + 3(1-x))
        g = int(round(x * N / (10 * x + 3 * (1 - x))))
        w = int(round((N - 10 * g) / 3))
        out.append((g, w))
    return out
