"""Device-resident engine on an MI355X: minimisation, MD segments, determinism, batch
invariance (a replica's trajectory is the same alone and inside a batch), precision modes
and the union-find cluster analysis."""
import numpy as np
import pytest

from molecular_simulation_code_amd import system as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _engine():
    from molecular_simulation_code_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    from molecular_simulation_code_amd import host
    return host


def _sys(ng, nw, seed, **kw):
    return S.make_system(ng, nw, seed=seed, params=S.Params(**kw))


def _engine_pairs(eng):
    """Per replica: float32 positions (device order), global ids and the list as a set of
    ordered (i, j) local index pairs."""
    import ctypes as C
    from molecular_simulation_code_amd._lib import lib, check, ptr
    N = eng.N
    a = np.empty((N, 4), np.float32)
    gid = np.empty(N, np.int32)
    check(lib().sofia_get_state(eng.ctx, ptr(a), None, ptr(gid)))
    na = C.c_int32()
    check(lib().sofia_get_nlist(eng.ctx, None, None, C.byref(na)))
    nl = np.empty(N * na.value, np.int32)
    nn = np.empty(N, np.int32)
    check(lib().sofia_get_nlist(eng.ctx, ptr(nl), ptr(nn), C.byref(na)))
    out, off = [], 0
    for n in eng.natoms:
        pairs = set()
        for k in range(n):
            row = nl[(off + k) * na.value:(off + k) * na.value + nn[off + k]]
            pairs.update((k, int(j) - 1) for j in row)
        out.append((a[off:off + n, :3].astype(np.float64), gid[off:off + n], pairs))
        off += n
    return out


def _pairs_within(x, L, r):
    from scipy.spatial import cKDTree
    t = cKDTree(np.mod(x, L), boxsize=L)
    p = t.query_pairs(r, output_type="ndarray")
    return {(int(i), int(j)) for i, j in p} | {(int(j), int(i)) for i, j in p}


def _excluded(s, gid):
    """Ordered local pairs excluded by specbond(1:num3bond), in the device order `gid`."""
    loc = np.empty(len(gid), np.int64)
    loc[gid - 1] = np.arange(len(gid))
    ex = set()
    for k, g in enumerate(gid):
        for e in s.specbond[g - 1, :s.num3bond[g - 1]]:
            ex.add((k, int(loc[e - 1])))
    return ex


def test_engine_list_holds_every_pair_inside_the_cutoff():
    """The engine's list (half-width cells, culled stencil rows) holds every non-excluded
    pair closer than rv when built, and every pair inside the force cutoff at any later
    step (complete-list rebuild criterion)."""
    host = _engine()
    systems = [S.make_system(4, 1500, seed=11), S.make_system(10, 3000, seed=12)]
    say("[engine list] np", [s.np for s in systems], "boxes", [float(s.box) for s in systems])
    e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=3) for s in systems],
                    precision=host.PREC_MIXED)
    try:
        e.init()
        d = systems[0].derived()
        rv = float(np.sqrt(d["rv2"]))
        rc = float(np.sqrt(max(d["rcut2"], d["cut_coulsq"])))
        for s, (x, gid, pairs) in zip(systems, _engine_pairs(e)):
            L = float(s.box)
            ex = _excluded(s, gid)
            must = _pairs_within(x, L, rv - 1e-3) - ex
            may = _pairs_within(x, L, rv + 1e-3) - ex
            assert must <= pairs, f"{len(must - pairs)} pairs closer than rv missing"
            assert pairs <= may, f"{len(pairs - may)} listed pairs beyond rv or excluded"
        e.minimize(30)
        e.run(80)
        assert e.rebuilds().min() >= 1
        for s, (x, gid, pairs) in zip(systems, _engine_pairs(e)):
            ex = _excluded(s, gid)
            need = _pairs_within(x, float(s.box), rc) - ex
            assert need <= pairs, f"{len(need - pairs)} pairs inside the cutoff not listed"
    finally:
        e.close()


def test_engine_list_complete_when_the_stencil_exceeds_one_lds_stage():
    """Cutoff 12.5 A with a 2 A skin: a cell's 125-cell stencil holds about 5,600 atoms,
    more than the 4,096 candidates one LDS stage takes, so every cell is built in two passes
    (the engine's answer to check_cache_size, Source/mod_neighbor.f90:313-388).  The lists
    must still hold every pair closer than rv, and the mixed pair forces must match the
    oracle's DSF on the same list and positions."""
    host = _engine()
    from oracle import oracle as O
    s = _sys(50, 1700, 31, rcut=12.5, cut_coul=12.5)
    d = s.derived()
    e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=5)], precision=host.PREC_MIXED)
    try:
        e.set_skin(2.0)
        e.init()
        rv = 12.5 + 2.0
        (x, gid, pairs), = _engine_pairs(e)
        L = float(s.box)
        ex = _excluded(s, gid)
        must = _pairs_within(x, L, rv - 1e-3) - ex
        may = _pairs_within(x, L, rv + 1e-3) - ex
        per_atom = len(must) / s.np
        say(f"[wide stencil] np {s.np}, box {L:.2f}, {per_atom:.0f} listed pairs per atom within rv")
        assert must <= pairs, f"{len(must - pairs)} pairs closer than rv missing"
        assert pairs <= may, f"{len(pairs - may)} listed pairs beyond rv or excluded"
        e.run(20)
        f, _ = e.pair_forces()
        xv, gidv = e.device_view()
        nl, nn, na = e.replica_list(0)
        prm = O.params_from(d, S.NTYPES)
        ref, _, _, _ = O.dsf_x64(xv, s.types[gidv - 1], s.q[gidv - 1], nl, nn, na, S.lj_tables(), prm)
        err = float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))
        say(f"[wide stencil] mixed pair forces vs oracle rel RMS {err:.2e}")
        assert err < 1e-5
    finally:
        e.close()


def test_engine_bonded_energies_match_oracle():
    """Bond, angle, torsion and improper energies of the engine (rigid waters evaluated inside the SHAKE
    kernel, the rest in the bonded kernels; torsions by the multiple-angle recurrence) against
    the oracle on the same positions, in a batch where the replicas' lists are concatenated."""
    host = _engine()
    from oracle import oracle as O
    systems = [S.make_system(5, 1000, seed=41), S.make_system(3, 1100, seed=42)]
    e = host.Engine(systems, precision=host.PREC_FP64)
    try:
        e.init()
        E = e.energies()
        for k, s in enumerate(systems):
            d = s.derived()
            kb = np.array([c[0] for c in S.BOND_COEFF]); r0 = np.array([c[1] for c in S.BOND_COEFF])
            ka = np.array([c[0] for c in S.ANGLE_COEFF]); th = np.radians([c[1] for c in S.ANGLE_COEFF])
            _, eb = O.bond_harmonic(s.bondlist, s.pos, s.types, kb, r0, d["box"], d["ibox"])
            _, ea = O.angle_harmonic(s.anglelist, s.pos, s.types, ka, th, d["box"], d["ibox"])
            tk = np.array([c[0] for c in S.DIHED_COEFF]); tm = np.array([c[1] for c in S.DIHED_COEFF], float)
            ts = np.array([c[2] for c in S.DIHED_COEFF], float)
            _, et = O.torsion_cos(host.torsion_list(s), s.pos, s.types, tk, tm, ts, d["box"], d["ibox"])
            ik = np.array([c[0] for c in S.IMPRO_K_CHI0]); ic = np.array([c[1] for c in S.IMPRO_K_CHI0])
            _, ei, _ = O.improper_harmonic(s.improlist, s.pos, s.types, ik, ic, d["box"], d["ibox"])
            say(f"[engine bonded] replica {k}: bond {E[k][4]:.6f} vs {eb:.6f}, angle {E[k][5]:.6f} vs {ea:.6f}, "
                f"torsion {E[k][6]:.6f} vs {et:.6f}, improper {E[k][8]:.6f} vs {ei:.6f}")
            np.testing.assert_allclose(E[k][6], et, rtol=2e-5, atol=1e-4)
            np.testing.assert_allclose(E[k][8], ei, rtol=2e-5, atol=1e-4)
            # the oracle follows the reference's real*4 coordinate differences; the engine
            # differences FP64 positions, so the sums agree to float32 rounding
            np.testing.assert_allclose(E[k][4], eb, rtol=2e-5, atol=1e-4)
            np.testing.assert_allclose(E[k][5], ea, rtol=2e-5, atol=1e-4)
    finally:
        e.close()


def test_engine_improper_forces_match_oracle():
    """The engine's improper forces, isolated as the difference of the initial forces of two
    engines on the same positions with and without the improper list, against the oracle's
    improper_harmonic_lammps14 on those positions.  The system has no SHAKE clusters: a
    constraint force depends on every force on its cluster, so it would not cancel.  The
    positions are jittered by 0.05 A: at the planar minimum chi = acos(c) sits at c = -1,
    where float32 and FP64 coordinate differences disagree far more than elsewhere."""
    import dataclasses
    host = _engine()
    from oracle import oracle as O
    s0 = S.make_system(12, 900, seed=43)
    jitter = np.random.default_rng(7).normal(0.0, 0.05, s0.pos.shape)
    s = dataclasses.replace(s0, pos=(s0.pos + jitter).astype(np.float32),
                            shake_num=np.zeros(0, np.int32),
                            shake_atoms=np.zeros((0, 4), np.int32),
                            shake_bond=np.zeros((0, 3), np.float64))
    bare = dataclasses.replace(s, improlist=np.zeros((0, 5), np.int32))
    f = []
    for sys_ in (s, bare):
        e = host.Engine([sys_], precision=host.PREC_FP64)
        try:
            e.init()
            f.append(e.forces()[0])
        finally:
            e.close()
    fi = f[0] - f[1]
    d = s.derived()
    ik = np.array([c[0] for c in S.IMPRO_K_CHI0]); ic = np.array([c[1] for c in S.IMPRO_K_CHI0])
    ref, _, _ = O.improper_harmonic(s.improlist, s.pos, s.types, ik, ic, d["box"], d["ibox"])
    rel = np.sqrt(((fi - ref) ** 2).sum() / (ref ** 2).sum())
    idle = np.setdiff1d(np.arange(s.np), s.improlist[:, :4].ravel() - 1)
    say(f"[engine improper] {len(s.improlist)} impropers: force rel RMS {rel:.2e}, "
        f"max |dF| on atoms in no improper {np.abs(fi[idle]).max():.3e}")
    assert (ref ** 2).sum() > 0
    # atoms in no improper get exactly nothing from it
    assert np.abs(fi[idle]).max() == 0.0
    assert rel < 1e-5


@pytest.mark.parametrize("term", ["bond", "angle", "torsion"])
def test_engine_bond_angle_torsion_forces_match_oracle(term):
    """Each of the engine's bond, angle and torsion forces, isolated as the difference of
    the initial forces of two engines on the same positions with and without that term's
    list (no constraints, so nothing else moves), against the oracle's bond_harmonic,
    angle_harmonic and the torsion series pinned to dihed_PME_amber."""
    import dataclasses
    host = _engine()
    from oracle import oracle as O
    s0 = S.make_system(12, 900, seed=44)
    jitter = np.random.default_rng(8).normal(0.0, 0.05, s0.pos.shape)
    s = dataclasses.replace(s0, pos=(s0.pos + jitter).astype(np.float32),
                            shake_num=np.zeros(0, np.int32),
                            shake_atoms=np.zeros((0, 4), np.int32),
                            shake_bond=np.zeros((0, 3), np.float64))
    d = s.derived()
    if term == "bond":
        bare = dataclasses.replace(s, bondlist=np.zeros((0, 3), np.int32))
        kb = np.array([c[0] for c in S.BOND_COEFF]); r0 = np.array([c[1] for c in S.BOND_COEFF])
        ref, _ = O.bond_harmonic(s.bondlist, s.pos, s.types, kb, r0, d["box"], d["ibox"])
        atoms = s.bondlist[:, :2]
    elif term == "angle":
        bare = dataclasses.replace(s, anglelist=np.zeros((0, 4), np.int32))
        ka = np.array([c[0] for c in S.ANGLE_COEFF]); th = np.radians([c[1] for c in S.ANGLE_COEFF])
        ref, _ = O.angle_harmonic(s.anglelist, s.pos, s.types, ka, th, d["box"], d["ibox"])
        atoms = s.anglelist[:, :3]
    else:
        bare = dataclasses.replace(s, n_gly=0)       # torsion_list() follows n_gly
        tk = np.array([c[0] for c in S.DIHED_COEFF]); tm = np.array([c[1] for c in S.DIHED_COEFF], float)
        ts = np.array([c[2] for c in S.DIHED_COEFF], float)
        tl = host.torsion_list(s)
        ref, _ = O.torsion_cos(tl, s.pos, s.types, tk, tm, ts, d["box"], d["ibox"])
        atoms = tl[:, :4]
    f = []
    for sys_ in (s, bare):
        e = host.Engine([sys_], precision=host.PREC_FP64)
        try:
            e.init()
            f.append(e.forces()[0])
        finally:
            e.close()
    fi = f[0] - f[1]
    rel = np.sqrt(((fi - ref) ** 2).sum() / (ref ** 2).sum())
    idle = np.setdiff1d(np.arange(s.np), atoms.ravel() - 1)
    say(f"[engine {term}] {len(atoms)} items: force rel RMS {rel:.2e}")
    assert (ref ** 2).sum() > 0
    if len(idle):                # atoms in no item of the term get exactly nothing from it
        assert np.abs(fi[idle]).max() == 0.0
    assert rel < 1e-5


def test_engine_cvff_impropers_match_oracle():
    """With Ewald electrostatics the engine evaluates the cvff impropers of the reference's
    PME wrappers (K = 1.1, d = -1, n = 2 from the topology): energy and isolated forces
    (with and without the improper list, no constraints) against the oracle."""
    import dataclasses
    host = _engine()
    from oracle import oracle as O
    s0 = S.make_system(12, 900, seed=44)
    jitter = np.random.default_rng(8).normal(0.0, 0.05, s0.pos.shape)
    s = dataclasses.replace(s0, pos=(s0.pos + jitter).astype(np.float32),
                            shake_num=np.zeros(0, np.int32), shake_atoms=np.zeros((0, 4), np.int32),
                            shake_bond=np.zeros((0, 3), np.float64))
    bare = dataclasses.replace(s, improlist=np.zeros((0, 5), np.int32))
    f, E = [], None
    for sys_ in (s, bare):
        e = host.Engine([sys_], precision=host.PREC_FP64, coul_flag=host.COUL_EWALD)
        try:
            e.init()
            f.append(e.forces()[0])
            if E is None:
                E = e.energies()[0]
        finally:
            e.close()
    d = s.derived()
    k = np.array([c[0] for c in S.IMPRO_COEFF]); sg = np.array([c[1] for c in S.IMPRO_COEFF], np.int32)
    mt = np.array([c[2] for c in S.IMPRO_COEFF], np.int32)
    ref, eref, _ = O.improper_cvff(s.improlist, s.pos, s.types, k, sg, mt, d["box"], d["ibox"])
    fi = f[0] - f[1]
    rel = np.sqrt(((fi - ref) ** 2).sum() / (ref ** 2).sum())
    say(f"[engine cvff] energy {E[8]:.6f} vs {eref:.6f}, force rel RMS {rel:.2e}")
    np.testing.assert_allclose(E[8], eref, rtol=2e-5, atol=1e-5)
    assert rel < 1e-5


def test_skin_changes_speed_not_trajectory():
    """Mixed precision: the skin (and lanes per atom) only change speed -- the trajectory is
    bit-identical because the lists always hold every pair inside the cutoff."""
    host = _engine()
    s = S.make_system(6, 1800, seed=21)
    outs = []
    for plan in ([], [(0, 0.6, 8), (30, 1.5, 32)]):
        say("[engine skin] plan", plan)
        e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=9)], precision=host.PREC_MIXED)
        try:
            e.init()
            done = 0
            for at, skin, lanes in plan + [(60, None, None)]:
                if at > done:
                    e.run(at - done)
                done = at
                if skin is not None:
                    e.set_skin(skin)
                    e.set_tuning(lanes, 256, 4)
            outs.append(e.state()[0])
        finally:
            e.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_bonded_bundles_match_per_kind_kernels():
    """The bonded items by molecule bundles (forces summed in LDS) and by per-kind kernels
    (global atomics) give the same integer sums: trajectories and energies agree bit for
    bit, also when the mode is switched mid-run, in both precisions."""
    host = _engine()
    specs = [host.ReplicaSpec(_sys(100, 600, 83), temp=300.0, gamma_ps=1.0, seed=93),
             host.ReplicaSpec(_sys(140, 900, 84), temp=310.0, gamma_ps=1.0, seed=94),
             host.ReplicaSpec(_sys(90, 700, 85), temp=290.0, gamma_ps=0.5, seed=95)]
    for prec in (host.PREC_MIXED, host.PREC_FP64):
        outs = {}
        for name, plan in (("kinds", [(0, 0)]), ("bundles", [(0, 1)]),
                           ("switch", [(0, 1), (15, 0), (30, 1)])):
            say(f"[engine bonded] precision {prec} plan {name}")
            e = host.Engine(specs, precision=prec)
            try:
                e.init()
                done = 0
                for at, mode in plan + [(45, None)]:
                    if at > done:
                        e.run(at - done)
                        done = at
                    if mode is not None:
                        e.set_bonded_mode(mode)
                outs[name] = (e.state(), e.energies(), e.forces())
            finally:
                e.close()
        ref = outs["kinds"]
        for name in ("bundles", "switch"):
            out = outs[name]
            for k in range(len(specs)):
                assert np.array_equal(out[0][k][0], ref[0][k][0]), f"{name}: replica {k} positions differ"
                assert np.array_equal(out[0][k][1], ref[0][k][1]), f"{name}: replica {k} velocities differ"
                assert np.array_equal(out[2][k], ref[2][k]), f"{name}: replica {k} forces differ"
            assert np.array_equal(out[1][:, 2:9], ref[1][:, 2:9]), f"{name}: bonded energies differ"


def test_charge_by_type_records_keep_the_trajectory():
    """Mixed precision: with charges set by atom type the per-atom pair kernels read 12-byte
    records and take C q_i q_j from a type-pair table; the compacting kernel (lanes 0) keeps
    16-byte records with the charges.  Both give the same trajectory bit for bit."""
    host = _engine()
    specs = [host.ReplicaSpec(_sys(100, 600, 86), temp=300.0, gamma_ps=1.0, seed=96),
             host.ReplicaSpec(_sys(110, 800, 87), temp=310.0, gamma_ps=1.0, seed=97)]
    outs = {}
    for lanes in (16, 0):
        say(f"[engine charge records] lanes {lanes}")
        e = host.Engine(specs, precision=host.PREC_MIXED)
        try:
            e.set_tuning(lanes, 256, 4)
            e.init()
            # the per-atom kernels really take the 12-byte records here
            assert e.config()["record_bytes"] == (12 if lanes else 16), e.config()
            e.minimize(20)
            e.run(40)
            outs[lanes] = (e.state(), e.forces())
        finally:
            e.close()
    for k in range(len(specs)):
        assert np.array_equal(outs[16][0][k][0], outs[0][0][k][0]), f"replica {k} positions differ"
        assert np.array_equal(outs[16][0][k][1], outs[0][0][k][1]), f"replica {k} velocities differ"
        assert np.array_equal(outs[16][1][k], outs[0][1][k]), f"replica {k} forces differ"


def test_minimize_then_nve_energy_is_stable():
    host = _engine()
    s = _sys(100, 600, 31)
    say("[engine] minimize + NVE, np", s.np)
    e = host.Engine([s], precision=host.PREC_FP64)
    try:
        e.init()
        e0 = e.energies()[0]
        e.minimize(300)
        e1 = e.energies()[0]
        assert e1[:7].sum() < e0[:7].sum()          # potential went down
        e.run(50)
        E = e.energies()[0]
        assert np.all(np.isfinite(E))
        tot1 = E.sum()
        e.run(50)
        tot2 = e.energies()[0].sum()
        say("  total energy", tot1, tot2)
        assert abs(tot2 - tot1) < 0.02 * abs(E[7]) + 5.0   # bounded drift over 50 steps
    finally:
        e.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_run_is_deterministic_and_graph_equals_eager(prec):
    host = _engine()
    P = host.PREC_FP64 if prec == "fp64" else host.PREC_MIXED
    s = _sys(100, 600, 32)
    outs = []
    for use_graph in (True, True, False):
        say(f"[engine {prec}] run graph={use_graph}")
        e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=5)], precision=P)
        try:
            e.init()
            e.minimize(100)
            e.run(40, use_graph=use_graph)
            x, v = e.state()[0]
            outs.append((x, v, e.energies()[0], e.rebuilds()[0]))
        finally:
            e.close()
    for k in (1, 2):
        assert np.array_equal(outs[0][0], outs[k][0])     # bit-identical positions
        assert np.array_equal(outs[0][1], outs[k][1])
        assert outs[0][3] == outs[k][3]


def test_replica_alone_equals_in_batch_mixed():
    host = _engine()
    a = host.ReplicaSpec(_sys(100, 600, 41), temp=300.0, gamma_ps=1.0, seed=11)
    b = host.ReplicaSpec(_sys(120, 900, 42), temp=290.0, gamma_ps=1.0, seed=12)
    c = host.ReplicaSpec(_sys(90, 700, 43), temp=310.0, gamma_ps=0.0, seed=13)

    def run(reps):
        e = host.Engine(reps, precision=host.PREC_MIXED)
        try:
            e.init()
            e.run(60)
            return e.state(), e.energies(), e.rebuilds()
        finally:
            e.close()

    say("[engine] alone vs batch")
    alone = [run([r]) for r in (a, b, c)]
    mix = run([a, b, c])
    for k in range(3):
        xa, va = alone[k][0][0]
        xm, vm = mix[0][k]
        assert np.array_equal(xa, xm), f"replica {k} positions differ"
        assert np.array_equal(va, vm), f"replica {k} velocities differ"
        assert alone[k][2][0] == mix[2][k]
        np.testing.assert_allclose(alone[k][1][0], mix[1][k], rtol=1e-12, atol=1e-9)


def test_mixed_tracks_fp64_over_short_run():
    host = _engine()
    s = _sys(100, 600, 51)
    res = {}
    for name, P in (("fp64", host.PREC_FP64), ("mixed", host.PREC_MIXED)):
        e = host.Engine([s], precision=P)
        try:
            e.init()
            e.minimize(200)
            x0, _ = e.state()[0]
            e.run(10, use_graph=False)
            res[name] = (x0, e.state()[0][0], e.energies()[0])
        finally:
            e.close()
    assert np.array_equal(res["fp64"][0] * 0, res["mixed"][0] * 0)
    dx = res["fp64"][1] - res["mixed"][1]
    L = float(s.box)
    dx -= L * np.round(dx / L)
    assert np.abs(dx).max() < 0.05          # trajectories stay together over 10 steps
    assert res["mixed"][2][:7].sum() == pytest.approx(res["fp64"][2][:7].sum(), rel=1e-3)


def test_cluster_union_find_matches_cpu_bfs():
    host = _engine()
    s = _sys(100, 600, 61)
    e = host.Engine([s], precision=host.PREC_MIXED)
    try:
        e.init()
        e.run(5, use_graph=False)
        big = e.largest_cluster(3.5)[0]
        x, _ = e.state()[0]
    finally:
        e.close()
    # CPU: molecules joined by an N(type 3)-O(types 4,5) contact < 3.5 A
    L = float(s.box)
    mol = s.mol
    ns = np.where((s.types == 3) & (s.moltype == 1))[0]
    os_ = np.where(np.isin(s.types, [4, 5]) & (s.moltype == 1))[0]
    adj = {m: set() for m in range(s.n_gly)}
    for i in ns:
        d = x[os_] - x[i]
        d -= L * np.round(d / L)
        close = os_[np.sqrt((d * d).sum(1)) < 3.5]
        for j in close:
            if mol[j] != mol[i]:
                adj[mol[i]].add(mol[j]); adj[mol[j]].add(mol[i])
    seen, best = set(), 0
    for m in range(s.n_gly):
        if m in seen:
            continue
        stack, size = [m], 0
        seen.add(m)
        while stack:
            u = stack.pop(); size += 1
            for w in adj[u]:
                if w not in seen:
                    seen.add(w); stack.append(w)
        best = max(best, size)
    assert big == best


def test_served_request_same_alone_and_in_mix():
    """Continuous batching: a request's result does not depend on what else is served."""
    host = _engine()
    from molecular_simulation_code_amd import server as SV
    pool = SV.request_pool([3000, 4200, 5200], minimize_steps=60)
    say("[serve] alone vs mix")

    def req(rid, k, arrival, seed, temp):
        s = pool[k].copy()
        s.v = S.maxwell_boltzmann(s, temp, np.random.default_rng(seed))
        return SV.Request(rid, arrival, s, seed, temp, nsteps=300)

    srv = SV.SegmentServer(segment=100, keep_state=True)
    alone, _, _ = srv.serve([req(0, 0, 0.0, 101, 300.0)], realtime=False)
    a = alone[0]
    mix_reqs = [req(0, 0, 0.0, 101, 300.0), req(1, 1, 0.0, 202, 285.0), req(2, 2, 1.5, 303, 315.0),
                req(3, 1, 2.5, 404, 295.0)]
    for policy in ("rebatch", "cohort"):
        say(f"[serve] policy {policy}")
        srv2 = SV.SegmentServer(segment=100, keep_state=True, policy=policy)
        mix, _, _ = srv2.serve(mix_reqs, realtime=False)
        if policy == "rebatch":
            assert srv2.rebatches >= 3                   # the batch really changed under request 0
        else:
            assert srv2.cohorts_started >= 3             # requests 2 and 3 started their own cohorts
        m = mix[0]
        assert np.array_equal(a.state[0], m.state[0]) and np.array_equal(a.state[1], m.state[1])
        assert a.clusters == m.clusters
        np.testing.assert_allclose(a.energies, m.energies, rtol=1e-12, atol=1e-9)
    assert all(r.finished >= r.arrival for r in mix)


def test_engine_restart_file_holds_the_replica_state(tmp_path):
    """A replica's state written as a restart file reads back as that state to the
    format's six decimals, in global-id order."""
    host = _engine()
    from molecular_simulation_code_amd import io as sio
    systems = [S.make_system(5, 1000, seed=51), S.make_system(3, 1100, seed=52)]
    e = host.Engine(systems, precision=host.PREC_MIXED)
    try:
        e.init()
        e.run(20)
        x, v = e.state()[1]
        p = tmp_path / "restart_20"
        e.write_restart(1, p, num=3, timeclock=0.04)
        box, temp, num, tc, xr, vr = sio.read_restart(p, systems[1].np)
        assert num == 3 and box == float(systems[1].box)
        np.testing.assert_allclose(xr, x, atol=5.1e-7 + 1e-5 * np.abs(x).max())
        np.testing.assert_allclose(vr, v, atol=5.1e-7)
    finally:
        e.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_rigid_waters_hold_their_geometry_exactly(prec):
    """Rigid waters go through SETTLE (closed form) instead of the iterative SHAKE, whose
    1e-4 tolerance leaves 1e-5 A errors: after MD steps every O-H and H-H distance holds to
    1e-8 A, and the step still conserves energy (checked elsewhere)."""
    host = _engine()
    s = S.make_system(4, 1100, seed=61)
    e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=3)],
                    precision=host.PREC_FP64 if prec == "fp64" else host.PREC_MIXED)
    try:
        e.init()
        e.run(60)
        x, _ = e.state()[0]
        w = s.shake_atoms[s.shake_num == -3] - 1
        bl = s.shake_bond[s.shake_num == -3]
        L = float(s.box)

        def dist(a, b):
            d = x[w[:, a]] - x[w[:, b]]
            d -= L * np.round(d / L)
            return np.sqrt((d * d).sum(1))
        err = max(np.abs(dist(0, 1) - bl[:, 0]).max(), np.abs(dist(0, 2) - bl[:, 1]).max(),
                  np.abs(dist(1, 2) - bl[:, 2]).max())
        say(f"[settle {prec}] {len(w)} waters, max constraint error {err:.2e} A")
        assert err < 1e-8
    finally:
        e.close()



def _lists_as_keys(eng, reps):
    """Per replica in `reps`: its Verlet list as sorted int64 keys i * n + j (local device
    order), its float32 positions, global ids and atom count."""
    import ctypes as C
    from molecular_simulation_code_amd._lib import lib, check, ptr
    N = eng.N
    a = np.empty((N, 4), np.float32)
    gid = np.empty(N, np.int32)
    check(lib().sofia_get_state(eng.ctx, ptr(a), None, ptr(gid)))
    na = C.c_int32()
    check(lib().sofia_get_nlist(eng.ctx, None, None, C.byref(na)))
    w = na.value
    nl = np.empty(N * w, np.int32)
    nn = np.empty(N, np.int32)
    check(lib().sofia_get_nlist(eng.ctx, ptr(nl), ptr(nn), C.byref(na)))
    out = []
    for rep in reps:
        off = int(np.sum(eng.natoms[:rep]))
        n = int(eng.natoms[rep])
        rows = nl[off * w:(off + n) * w].reshape(n, w)
        cnt = nn[off:off + n].astype(np.int64)
        assert cnt.max() <= w
        i = np.repeat(np.arange(n, dtype=np.int64), cnt)
        j = rows[np.arange(w)[None, :] < cnt[:, None]].astype(np.int64) - 1
        assert j.min() >= 0 and j.max() < n
        out.append((np.sort(i * n + j), a[off:off + n, :3].astype(np.float64),
                    gid[off:off + n], n))
    return out


def _within_keys(x, L, r, n):
    from scipy.spatial import cKDTree
    p = cKDTree(np.mod(x, L), boxsize=L).query_pairs(r, output_type="ndarray").astype(np.int64)
    return np.sort(np.concatenate([p[:, 0] * n + p[:, 1], p[:, 1] * n + p[:, 0]]))


def _missing(a, b):
    """Boolean mask over sorted keys a: True where the key is absent from sorted keys b."""
    k = np.minimum(np.searchsorted(b, a), max(len(b) - 1, 0))
    return ~((len(b) > 0) & (b[k] == a))


def _check_full_list(keys, x, gid, n, s, rv, L):
    """Verlet list of one replica against the periodic pair set, with a 1e-3 A band at rv
    for float32 rounding: no duplicates, no excluded pair, every non-excluded pair closer
    than rv - 1e-3 listed, nothing at rv + 1e-3 or beyond; (i, j) listed without (j, i)
    only inside the band."""
    assert len(np.unique(keys)) == len(keys), "duplicate list entries"
    ex = np.array(sorted(a * n + b for a, b in _excluded(s, gid)), np.int64)
    assert not (~_missing(ex, keys)).any(), "excluded pairs listed"
    must = _within_keys(x, L, rv - 1e-3, n)
    must = must[_missing(must, ex)]
    assert not _missing(must, keys).any(), "pairs closer than rv missing"
    assert not _missing(keys, _within_keys(x, L, rv + 1e-3, n)).any(), "pairs beyond rv listed"
    tkeys = np.sort((keys % n) * n + keys // n)
    one_way = keys[_missing(keys, tkeys)]
    d = x[one_way // n] - x[one_way % n]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(axis=1))
    assert (np.abs(r - rv) < 1e-3).all(), "one-way list entries away from rv"
    return len(one_way)


def test_full_size_batch_properties():
    """1.27 M atoms in one mixed-precision batch: 23 copies of the 64-replica workload's
    largest replica (55,344 atoms).  The first and last copy's Verlet lists hold every
    non-excluded pair closer than rv, nothing beyond it and are symmetric up to float32
    rounding at rv; copies with the same positions get bit-identical forces wherever they
    sit in the batch, before and after minimisation, whatever their velocities; and every
    replica's total force sums to zero up to the float32 rounding of the pair terms."""
    host = _engine()
    comps = S.replica_sizes(64)
    k = int(np.argmax([10 * g + 3 * w for g, w in comps]))
    s = S.make_system(*comps[k], seed=20261210 + k)
    ncopy = 23
    say("[full size] replica", k, "np", s.np, "copies", ncopy, "atoms", ncopy * s.np)
    # NVE (gamma 0): the stored force is then the potential's alone, with no Langevin draw,
    # so copies with different velocity seeds must still agree bit for bit
    specs = [host.ReplicaSpec(s, temp=300.0, gamma_ps=0.0, seed=500 + c) for c in range(ncopy)]
    e = host.Engine(specs, precision=host.PREC_MIXED)
    try:
        e.init()
        d = s.derived()
        rv = float(np.sqrt(d["rv2"]))
        for rep, (keys, x, gid, n) in zip((0, ncopy - 1), _lists_as_keys(e, (0, ncopy - 1))):
            one_way = _check_full_list(keys, x, gid, n, s, rv, float(s.box))
            say(f"[full size] list of copy {rep}: {len(keys)} entries, {one_way} one-way at rv")
        for stage in ("init", "minimize"):
            if stage == "minimize":
                e.minimize(20)
            f = e.forces()
            say(f"[full size] {stage}: max |F| {np.abs(f[0]).max():.3e}")
            assert np.abs(f[0]).max() > 1.0
            for c in range(1, ncopy):
                assert np.array_equal(f[c], f[0]), f"{stage}: copy {c} forces differ from copy 0"
            for c in (0, ncopy - 1):
                net = np.abs(f[c].sum(axis=0)).max()
                scale = np.abs(f[c]).sum()
                say(f"[full size] {stage} copy {c}: |sum F| {net:.3e}, sum |F| {scale:.3e}")
                # measured 8e-11 of sum |F| (init) and 2.5e-11 (minimised): the pair terms of
                # (i, j) and (j, i) differ only in the float32 rounding of q_i q_j
                assert net <= 1e-9 * scale
    finally:
        e.close()


def test_full_size_mixed_forces_and_energies_track_fp64():
    """The same 1.27 M-atom batch (23 copies of the 64-replica workload's largest replica) in
    mixed precision and in FP64, NVE, at the generated positions: per replica the mixed forces
    are within 1e-5 relative RMS of the FP64 forces and the potential energy within 1e-6."""
    host = _engine()
    comps = S.replica_sizes(64)
    k = int(np.argmax([10 * g + 3 * w for g, w in comps]))
    s = S.make_system(*comps[k], seed=20261210 + k)
    ncopy = 23
    specs = [host.ReplicaSpec(s, temp=300.0, gamma_ps=0.0, seed=700 + c) for c in range(ncopy)]
    out = {}
    for name, prec in (("fp64", host.PREC_FP64), ("mixed", host.PREC_MIXED)):
        say("[full size precision]", name, "atoms", ncopy * s.np)
        e = host.Engine(specs, precision=prec)
        try:
            e.init()
            out[name] = (e.forces(), e.energies())
        finally:
            e.close()
    (fd, ed), (fm, em) = out["fp64"], out["mixed"]
    pot = [0, 1, 2, 3, 4, 5, 6, 8]          # every potential term; column 7 is the kinetic energy
    for c in (0, ncopy - 1):
        rms = np.sqrt(((fm[c] - fd[c]) ** 2).sum() / (fd[c] ** 2).sum())
        ud, um = ed[c, pot].sum(), em[c, pot].sum()
        say(f"[full size precision] copy {c}: force rel RMS {rms:.3e}, potential "
            f"{ud:.6f} vs {um:.6f} (rel {abs(um - ud) / abs(ud):.2e}), terms rel "
            + " ".join(f"{abs(em[c, t] - ed[c, t]) / max(abs(ed[c, t]), 1e-30):.1e}" for t in pot))
        assert rms <= 1e-5
        assert abs(um - ud) <= 1e-6 * abs(ud)


def test_settle_failure_is_reported():
    """A rigid water whose unconstrained step leaves SETTLE's reach (one H launched at
    2 A/fs) makes the MD call fail with SOFIA_E_SHAKE instead of writing NaN positions."""
    host = _engine()
    from molecular_simulation_code_amd._lib import SofiaError
    s = _sys(4, 1100, 71)
    h = 10 * s.n_gly + 1                      # first water's first H (0-based)
    s.v[h] = np.array([0.0, 0.0, 2.0], np.float32)
    e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=0.0, seed=1)], precision=host.PREC_FP64)
    try:
        with pytest.raises(SofiaError) as ei:
            e.init()
            e.run(2)
        say("[settle failure]", ei.value)
        assert ei.value.code == -6
    finally:
        e.close()


def test_forcefield_change_mid_run_graph_equals_eager():
    """New LJ coefficients uploaded into a running mixed-precision engine (charges by type,
    12-byte records and a type-pair table) reach the captured step: graph and eager runs
    stay bit-identical after the change, and differ from a run without it."""
    host = _engine()
    import ctypes as C
    from molecular_simulation_code_amd._lib import check, lib, ptr
    s = _sys(40, 1000, 72)
    lj = S.lj_tables()
    mod = [np.ascontiguousarray(x * np.float32(1.05) if k < 2 else x, dtype=np.float32)
           for k, x in enumerate(lj[:4])]
    l14 = [np.ascontiguousarray(x, np.float64) for x in lj[4:]]
    mass = np.ascontiguousarray(S.type_mass(), np.float64)
    outs = {}
    for tag, use_graph, change in (("graph", True, True), ("eager", False, True),
                                   ("unchanged", True, False)):
        e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=9)],
                        precision=host.PREC_MIXED)
        try:
            e.init()
            e.minimize(20)
            e.run(20, use_graph=use_graph)
            assert e.config()["record_bytes"] == 12
            if change:
                check(lib().sofia_set_forcefield(e.ctx, *[ptr(a) for a in mod + l14 + [mass]]))
                assert e.config()["record_bytes"] == 12   # the pair table was rebuilt
            e.run(20, use_graph=use_graph)
            outs[tag] = e.state()[0]
        finally:
            e.close()
    assert np.array_equal(outs["graph"][0], outs["eager"][0])
    assert np.array_equal(outs["graph"][1], outs["eager"][1])
    assert not np.array_equal(outs["graph"][0], outs["unchanged"][0])


def test_list_overflow_mid_segment_is_grown_and_redone():
    """With no headroom in the list stride (headroom 1.0: the stride is the largest count at
    set-up), a rebuild inside the segment overflows; the segment is run again from its
    starting state with a grown stride and ends bit-identical to a run with ample room,
    in one segment and across several, with multi-step graphs."""
    host = _engine()
    # hot replicas (600-650 K) rebuild often and their largest neighbour count moves
    s1, s2 = _sys(20, 1400, 81, temp=600.0), _sys(30, 1700, 82, temp=650.0)
    specs = [host.ReplicaSpec(s1, temp=600.0, gamma_ps=1.0, seed=21),
             host.ReplicaSpec(s2, temp=650.0, gamma_ps=1.0, seed=22)]
    outs = {}
    for tag, headroom, segs, gsteps in (("ample", 1.5, [200], 1), ("tight", 1.0, [200], 1),
                                        ("tight-segments", 1.0, [70, 70, 60], 8)):
        e = host.Engine(specs, precision=host.PREC_MIXED)
        try:
            e.set_list_headroom(headroom)
            e.set_graph_steps(gsteps)
            e.init()
            na0, _ = e.list_stats()
            for n in segs:
                e.run(n)
            na, redos = e.list_stats()
            say(f"[list redo] {tag}: stride {na0} -> {na}, {redos} segments redone, "
                f"rebuilds {e.rebuilds().tolist()}")
            outs[tag] = (e.state(), e.forces(), e.energies(), redos, na0, na)
        finally:
            e.close()
    assert outs["ample"][3] == 0
    assert outs["tight"][3] >= 1 and outs["tight"][5] > outs["tight"][4]
    assert outs["tight-segments"][3] >= 1
    for tag in ("tight", "tight-segments"):
        for k in range(len(specs)):
            assert np.array_equal(outs[tag][0][k][0], outs["ample"][0][k][0]), f"{tag} {k} x"
            assert np.array_equal(outs[tag][0][k][1], outs["ample"][0][k][1]), f"{tag} {k} v"
            assert np.array_equal(outs[tag][1][k], outs["ample"][1][k]), f"{tag} {k} F"
        np.testing.assert_allclose(outs[tag][2], outs["ample"][2], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("gsteps", [10, 50])
def test_multi_step_graphs_equal_one_step_graphs(gsteps):
    """Replaying k-step graphs (and one-step graphs for the rest of a segment) gives the
    one-step-graph trajectory bit for bit."""
    host = _engine()
    s = _sys(60, 1500, 83)
    outs = []
    for k in (1, gsteps):
        e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=31)],
                        precision=host.PREC_MIXED)
        try:
            e.set_graph_steps(k)
            e.init()
            e.run(73)
            e.run(gsteps + 5)
            outs.append((e.state()[0], e.forces()[0], e.rebuilds()[0]))
        finally:
            e.close()
    assert np.array_equal(outs[0][0][0], outs[1][0][0]) and np.array_equal(outs[0][0][1], outs[1][0][1])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


def test_greatest_cluster_matches_oracle_exactly():
    """get_greatestCluster on the device (molecule centres, density counts, union-find over
    dense pairs, first largest cluster plus its non-dense neighbours) against the oracle's
    restatement on the engine's own FP64 positions: the same integers for every replica and
    several cuts and density thresholds, before and after MD steps."""
    host = _engine()
    from oracle import oracle as O
    systems = [S.make_system(220, 1000, seed=91), S.make_system(150, 1400, seed=92)]
    e = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=1.0, seed=41 + k)
                     for k, s in enumerate(systems)], precision=host.PREC_MIXED)
    try:
        e.init()
        for stage in ("init", "after 40 steps"):
            if stage != "init":
                e.run(40)
            x, gid = e.device_view()
            off = np.concatenate([[0], np.cumsum(e.natoms)])
            for cut, dmin in ((6.0, 2), (7.5, 4), (9.0, 6), (5.0, 1)):
                tot, core, surf = e.greatest_cluster(cut, dmin)
                for k, s in enumerate(systems):
                    sl = slice(int(off[k]), int(off[k + 1]))
                    g = gid[sl]
                    mol = np.where(s.moltype[g - 1] == 1, s.mol[g - 1], -1).astype(np.int32)
                    com = O.molecule_centers(x[sl], g, mol, s.n_gly, s.box)
                    t2, c2, s2, _, _ = O.greatest_cluster(com, s.box, cut, dmin)
                    say(f"[greatest cluster {stage}] replica {k} cut {cut} densmin {dmin}: "
                        f"device {tot[k]} ({core[k]} + {surf[k]}), oracle {t2} ({c2} + {s2})")
                    assert (tot[k], core[k], surf[k]) == (t2, c2, s2)
            assert core.max() > 1
    finally:
        e.close()
