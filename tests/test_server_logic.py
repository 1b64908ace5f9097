"""Host logic on the CPU.  The serving loop: admission by HBM footprint, continuous batching
(cohorts that form at segment boundaries, finish, re-batch their survivors and merge past
max_cohorts), and complete results; and the autotuner's cache of tunings.  The GPU engine is replaced by a stand-in that records
what the server asks of it, so these tests need no device; the GPU tests
(tests/test_gpu_engine.py, tests/test_gpu_configs.py) run the same loop on the real engine."""
import numpy as np
import pytest

from molecular_simulation_code_amd import server as SV
from molecular_simulation_code_amd._lib import NENERGY


class _Tiny:
    """Only what the server reads of a system: np, pos, v."""

    def __init__(self, n):
        self.np = n
        self.pos = np.zeros((n, 3), np.float32)
        self.v = np.zeros((n, 3), np.float32)


class _FakeEngine:
    """Stand-in for host.Engine: tracks the atoms resident at once and the steps run."""
    resident = 0
    peak = 0
    built = 0

    def __init__(self, reps, precision=None):
        self.reps = reps
        self.n = len(reps)
        self._steps = np.zeros(self.n, np.int64)
        self.atoms = sum(r.system.np for r in reps)
        _FakeEngine.resident += self.atoms
        _FakeEngine.peak = max(_FakeEngine.peak, _FakeEngine.resident)
        _FakeEngine.built += 1

    def put_state(self, xv):
        assert len(xv) == self.n

    def set_steps(self, steps):
        self._steps[:] = steps

    def resume(self, forces):
        assert len(forces) == self.n

    def launch(self, n):
        self._steps += n

    def run(self, n):
        self._steps += n

    def finish(self):
        pass

    def largest_cluster(self, rc):
        return np.arange(self.n)

    def energies(self):
        return np.tile(np.arange(NENERGY, dtype=float), (self.n, 1))

    def state(self):
        return [(r.system.pos.astype(np.float64), r.system.v.astype(np.float64)) for r in self.reps]

    def forces(self):
        return [np.zeros((r.system.np, 3)) for r in self.reps]

    def steps(self):
        return self._steps.copy()

    def close(self):
        _FakeEngine.resident -= self.atoms
        self.atoms = 0


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(SV, "Engine", _FakeEngine)
    _FakeEngine.resident = _FakeEngine.peak = _FakeEngine.built = 0
    return _FakeEngine


def _stream(sizes, arrivals, nsteps):
    return [SV.Request(rid=k, arrival=t, system=_Tiny(n), seed=k + 1, temp=300.0, nsteps=s)
            for k, (n, t, s) in enumerate(zip(sizes, arrivals, nsteps))]


@pytest.mark.parametrize("policy", ["cohort", "rebatch"])
def test_admission_never_exceeds_the_hbm_budget(fake, policy):
    """Requests whose footprint does not fit wait until earlier ones finish; the atoms
    resident on the (stand-in) device never exceed budget / bytes_per_atom."""
    bpa = 6000
    budget_atoms = 30000
    srv = SV.SegmentServer(hbm_budget_gb=budget_atoms * bpa / 1e9, segment=100,
                           bytes_per_atom=bpa, policy=policy)
    srv._calibrated = True                    # no device to measure bytes per atom on
    reqs = _stream([12000, 15000, 9000, 20000, 3000, 25000],
                   [0.0, 0.0, 0.5, 1.0, 1.0, 2.0], [300, 500, 200, 400, 100, 300])
    res, wall, atom_steps = srv.serve(reqs, realtime=False)
    assert [r.rid for r in res] == list(range(len(reqs)))
    assert fake.peak <= budget_atoms, (fake.peak, budget_atoms)
    assert fake.resident == 0                 # every engine closed
    assert atom_steps == sum(r.system.np * r.nsteps for r in reqs)
    for r, q in zip(res, reqs):
        assert r.admitted >= q.arrival and r.finished > r.admitted
        assert len(r.clusters) == -(-q.nsteps // 100)   # one per segment it ran


def test_cohorts_form_at_segment_boundaries_and_merge(fake):
    """Arrivals at different segment boundaries start cohorts of their own, which run side by
    side; past max_cohorts they are merged into one engine."""
    srv = SV.SegmentServer(hbm_budget_gb=100.0, segment=50, bytes_per_atom=6000,
                           policy="cohort", max_cohorts=2)
    srv._calibrated = True
    reqs = _stream([1000, 2000, 3000, 4000], [0.0, 1.0, 2.0, 3.0], [400, 400, 400, 400])
    res, wall, _ = srv.serve(reqs, realtime=False)
    assert len(res) == 4 and fake.resident == 0
    assert srv.cohorts_started == 4           # each boundary's arrivals formed a cohort
    assert srv.rebatches >= 1                 # a third cohort exceeded max_cohorts: merged
    for r in res:
        assert r.admitted == pytest.approx(reqs[r.rid].arrival)   # admitted on arrival


def test_a_request_waits_until_the_resident_footprint_leaves_room(fake):
    """Two requests that each fill the whole budget: the second is admitted only when the
    first has finished and left the device."""
    bpa = 6000
    srv = SV.SegmentServer(hbm_budget_gb=10000 * bpa / 1e9, segment=100, bytes_per_atom=bpa)
    srv._calibrated = True
    reqs = _stream([10000, 10000], [0.0, 0.0], [100, 100])
    res, _, _ = srv.serve(reqs, realtime=False)
    assert len(res) == 2
    assert fake.peak == 10000                 # the second waited for the first to finish
    assert res[1].admitted >= res[0].finished


def test_autotune_cache_survives_the_process_and_buckets_by_size_and_density(monkeypatch, tmp_path):
    """The autotuner's cache: a bucket written by one process (SOFIA_TUNE_CACHE names a JSON
    file) is read back by the next; the key separates replica count, size and density on a
    coarse scale and is shared within a bucket."""
    from molecular_simulation_code_amd import host as H
    path = tmp_path / "tune.json"
    monkeypatch.setenv("SOFIA_TUNE_CACHE", str(path))
    monkeypatch.setattr(H, "_TUNE_CACHE", {})
    H._tune_cache_put("k1", (1.6, 16, 2.4, 16, 1, 1))
    monkeypatch.setattr(H, "_TUNE_CACHE", {})          # a fresh process
    assert H._tune_cache_get("k1") == (1.6, 16, 2.4, 16, 1, 1)
    assert H._tune_cache_get("k2") is None
    path.write_text("not json")                         # a damaged file is ignored
    monkeypatch.setattr(H, "_TUNE_CACHE", {})
    assert H._tune_cache_get("k1") is None

    class _P:
        precision, coul_flag = 1, 0

    def key(natoms, box, nrep=1):
        e = H.Engine.__new__(H.Engine)
        e.params = _P()
        e.reps = [H.ReplicaSpec(type("Sys", (), {"box": box})()) for _ in range(nrep)]
        e.N = natoms * nrep
        return e.tune_key()

    assert key(204000, 127.3) == key(205000, 127.5)     # same bucket
    assert key(204000, 127.3) != key(1020000, 217.7)    # another size
    assert key(204000, 127.3) != key(204000, 127.3, nrep=4)
    assert key(204000, 127.3) != key(204000, 140.0)     # another density
