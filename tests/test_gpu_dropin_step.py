"""The lean drop-in step (host.DropinStep: page-locked arrays, each routine copying only the
arrays it reads and writes) against the same step through SofiaGPU (every routine copying
the whole state): the reference's DSF time loop (Source/mod_integrate.f90:270-308) must give
the same trajectory bit for bit, list rebuilds included."""
import numpy as np
import pytest

from molecular_simulation_code_amd import system as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _full_step(g, dthalf):
    """The step as SofiaGPU's routines run it (the earlier drop-in harness)."""
    g.nve_v(dthalf)
    g.nve_x()
    g.pbc()
    rebuilt = bool(g.check_neighbor())
    if rebuilt:
        g.neighbor_wrapper()
    f = g.lj_cut_coul_dsf_nonewton(0).astype(np.float64)
    f += g.pair14_dsf()
    f += g.bonded_forces()
    g.ff = f.astype(np.float32)
    g.shake(1)
    g.nve_v(dthalf)
    return rebuilt


def test_lean_dropin_step_matches_the_full_state_step():
    from molecular_simulation_code_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    from molecular_simulation_code_amd.host import DropinStep, SofiaGPU
    params = S.Params(rcut=5.5, cut_coul=5.5, pad=1.0)
    s = S.make_system(20, 400, seed=41, params=params)
    ga, gb = SofiaGPU(s.copy()), SofiaGPU(s.copy())
    d = None
    try:
        dthalf = float(s.derived()["dthalf"])
        for g in (ga, gb):
            g.neighbor_wrapper(0)
        f = ga.lj_cut_coul_dsf_nonewton(0).astype(np.float64)
        f += ga.pair14_dsf()
        f += ga.bonded_forces()
        ga.ff = f.astype(np.float32)
        ga.shake(1)
        d = DropinStep(gb)
        d.forces()
        assert np.array_equal(d.ff, ga.ff)
        builds = [0, 0]
        for step in range(40):
            builds[0] += _full_step(ga, dthalf)
            builds[1] += d.step()
            pa = np.concatenate([ga.position, ga.type.view(np.float32)[:, None]], axis=1)
            assert np.array_equal(d.pos, pa), f"positions differ at step {step}"
            assert np.array_equal(d.v, ga.v), f"velocities differ at step {step}"
            assert np.array_equal(d.ff, ga.ff), f"forces differ at step {step}"
        say(f"[drop-in step] 40 steps of {s.np} atoms bit-identical; list builds {builds}")
        assert builds[0] == builds[1] and builds[0] >= 1
    finally:
        if d is not None:
            d.close()
        ga.close()
        gb.close()


def _device_topology(ctx, n, nb, na, nd, ni, ns):
    """sofia_get_topology into fresh host arrays (the host's layout: 1-based ids, 0 = empty)."""
    from molecular_simulation_code_amd._lib import check, lib, ptr
    i32 = np.int32
    T = dict(q=np.empty(n, np.float32), num1bond=np.empty(n, i32), num2bond=np.empty(n, i32),
             num3bond=np.empty(n, i32), specbond=np.empty((n, S.SPEC_SLOTS), i32),
             bondlist=np.empty((nb, 3), i32), anglelist=np.empty((na, 4), i32),
             dihedlist=np.empty((nd, 5), i32), improlist=np.empty((ni, 5), i32),
             shake_atoms=np.empty((ns, 4), i32))
    check(lib().sofia_get_topology(ctx, *[ptr(T[k]) for k in (
        "q", "num1bond", "num2bond", "num3bond", "specbond", "bondlist", "anglelist",
        "dihedlist", "improlist", "shake_atoms")]))
    return T


def _check_by_global_id(T, gid, s, tag):
    """Every returned array, read through the device order's global ids, is the system's
    original topology: the permutation moved rows and renamed atoms, nothing else."""
    o = gid - 1                                       # original index of each current atom
    for k in ("q", "num1bond", "num2bond", "num3bond"):
        assert np.array_equal(T[k], getattr(s, k)[o]), f"{tag}: {k}"
    g0 = np.concatenate([[0], gid]).astype(np.int32)  # current 1-based id -> original id
    sp = g0[T["specbond"]]
    want = s.specbond[o]
    used = np.arange(S.SPEC_SLOTS)[None, :] < s.num3bond[o][:, None]
    assert np.array_equal(np.where(used, sp, 0), np.where(used, want, 0)), f"{tag}: specbond"
    for k, nidx in (("bondlist", 2), ("anglelist", 3), ("dihedlist", 4), ("improlist", 4),
                    ("shake_atoms", 4)):
        got = T[k].copy()
        got[:, :nidx] = g0[got[:, :nidx]]
        assert np.array_equal(got, getattr(s, k)), f"{tag}: {k}"


def test_device_topology_equals_the_host_reindex():
    """After list builds at several states, sofia_get_topology returns what the host's own
    re-index by tt gives (sort_specbond_numbond and the sort_* list routines,
    Source/mod_neighbor.f90:1011-1168), and read through global ids it is the original
    topology."""
    from molecular_simulation_code_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    from molecular_simulation_code_amd.host import SofiaGPU
    params = S.Params(rcut=5.5, cut_coul=5.5, pad=1.0)
    s = S.make_system(20, 400, seed=43, params=params)
    g = SofiaGPU(s.copy())
    try:
        dthalf = float(s.derived()["dthalf"])
        imp = s.improlist.copy()          # the host's improlist, re-indexed at every build
        host_nw = g.neighbor_wrapper

        def neighbor_wrapper(step=0):
            tt = host_nw(step)
            imp[:, :4] = np.concatenate([[0], tt]).astype(np.int32)[imp[:, :4]]
            return tt

        g.neighbor_wrapper = neighbor_wrapper   # also the builds the MD steps trigger
        for cycle in range(4):
            tt = g.neighbor_wrapper(0)
            T = _device_topology(g.ctx, g.np, len(g.bondlist), len(g.anglelist),
                                 len(g.dihedlist), len(imp), len(g.shake_atoms))
            for k in ("q", "num1bond", "num2bond", "num3bond", "specbond", "bondlist",
                      "anglelist", "dihedlist", "shake_atoms"):
                assert np.array_equal(T[k], getattr(g, k)), f"build {cycle}: {k}"
            assert np.array_equal(T["improlist"], imp), f"build {cycle}: improlist"
            _check_by_global_id(T, g.global_id, s, f"build {cycle}")
            moved = int(np.count_nonzero(tt != np.arange(1, g.np + 1)))
            say(f"[topology] build {cycle}: {moved} of {g.np} atoms moved; arrays equal")
            for _ in range(15):
                _full_step(g, dthalf)
    finally:
        g.close()


def test_device_topology_of_a_batch_by_global_id():
    """A batch of three replicas after an MD segment with list builds: each replica's part of
    every returned array uses that replica's own 1-based ids and, read through its global
    ids, is its original topology."""
    from molecular_simulation_code_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    from molecular_simulation_code_amd import host
    from molecular_simulation_code_amd._lib import check, lib, ptr
    systems = [S.make_system(ng, nw, seed=sd) for ng, nw, sd in ((100, 600, 61), (120, 900, 62),
                                                                  (90, 700, 63))]
    e = host.Engine([host.ReplicaSpec(x, temp=300.0, gamma_ps=1.0, seed=k + 1)
                     for k, x in enumerate(systems)], precision=host.PREC_MIXED)
    try:
        e.init()
        e.run(100)
        builds = e.rebuilds()
        assert min(builds) >= 1, builds
        cnt = {k: [len(getattr(x, k)) for x in systems]
               for k in ("bondlist", "anglelist", "dihedlist", "improlist", "shake_atoms")}
        T = _device_topology(e.ctx, e.N, *[sum(cnt[k]) for k in (
            "bondlist", "anglelist", "dihedlist", "improlist", "shake_atoms")])
        gid = np.empty(e.N, np.int32)
        check(lib().sofia_get_state(e.ctx, None, None, ptr(gid)))
        a0 = 0
        item0 = {k: 0 for k in cnt}
        for r, x in enumerate(systems):
            n = x.np
            part = {k: T[k][a0:a0 + n] for k in ("q", "num1bond", "num2bond", "num3bond",
                                                  "specbond")}
            for k in cnt:
                part[k] = T[k][item0[k]:item0[k] + cnt[k][r]]
                item0[k] += cnt[k][r]
            _check_by_global_id(part, gid[a0:a0 + n], x, f"replica {r}")
            a0 += n
        say(f"[topology] batch of 3 after {list(builds)} list builds: every replica's arrays "
            "equal its original topology")
    finally:
        e.close()


def test_force_wrapper_equals_the_three_routines_and_the_host_merge():
    """sofia_force_wrapper_dsf (force_wrapper_DSF_NOMIC, Source/mod_force.f90:508-551) gives,
    bit for bit, the forces of sofia_lj_cut_coul_dsf + sofia_pair14_dsf + sofia_bonded_forces
    merged on the host in FP64 and stored real*4, and the same energies and virial."""
    from molecular_simulation_code_amd import _lib
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    from molecular_simulation_code_amd._lib import check, lib, ptr
    from molecular_simulation_code_amd.host import SofiaGPU
    params = S.Params(rcut=5.5, cut_coul=5.5, pad=1.0)
    s = S.make_system(20, 400, seed=45, params=params)
    g = SofiaGPU(s)
    try:
        g.neighbor_wrapper(0)
        n = g.np
        one = lambda: np.zeros(1)
        parts = {k: np.empty((n, 3), np.float32) for k in ("pair", "p14", "bonded")}
        e = [one() for _ in range(8)]
        vir = np.zeros(3)
        check(lib().sofia_lj_cut_coul_dsf(g.ctx, ptr(parts["pair"]), None, ptr(e[0]), ptr(e[1]),
                                          ptr(vir)))
        check(lib().sofia_pair14_dsf(g.ctx, ptr(parts["p14"]), ptr(e[2]), ptr(e[3])))
        check(lib().sofia_bonded_forces(g.ctx, ptr(parts["bonded"]), *[ptr(x) for x in e[4:]]))
        merged = ((parts["pair"].astype(np.float64) + parts["p14"]) + parts["bonded"]).astype(np.float32)
        ff = np.empty((n, 3), np.float32)
        ew = np.zeros(8)
        virw = np.zeros(3)
        check(lib().sofia_force_wrapper_dsf(g.ctx, ptr(ff), ptr(ew), ptr(virw)))
        assert np.array_equal(ff, merged)
        assert np.array_equal(ew, np.concatenate(e))
        assert np.array_equal(virw, vir)
        assert np.count_nonzero(parts["p14"]) and np.count_nonzero(parts["bonded"])
        say(f"[force wrapper] {n} atoms: forces, 8 energies and virial equal the three routines")
    finally:
        g.close()
