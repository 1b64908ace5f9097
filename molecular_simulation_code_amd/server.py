"""Continuous-batching server for MD-segment requests.

Requests arrive over time, each carrying a system (coordinates and topology), a seed, a
temperature and a step count.  The server keeps every admitted request's state (FP64
positions and velocities, last force, Langevin step counter) resident on the GPU and
advances them ``segment`` steps at a time.  At each segment boundary finished requests leave
and waiting ones are admitted while their HBM footprint fits the budget -- the set of
requests on the GPU changes as work arrives (continuous batching).

Two policies:

* ``cohort`` (default): the requests admitted at one segment boundary form a batched engine
  of their own; every segment, all cohorts' steps are enqueued back to back on their own
  streams and run concurrently, so admitting work never rebuilds the batch of the requests
  already running.  A cohort whose requests partly finish is re-batched with the survivors;
  more than ``max_cohorts`` cohorts are merged into one.
* ``rebatch``: one engine holds every active request and is rebuilt whenever the set
  changes.

A replica that moves into a new batch is resumed with its exact FP64 state, last total
force and step counter; with the mixed-precision engine (order-independent int64 force
sums, counter-based Langevin noise, lists that hold every pair inside the cutoff) its
trajectory is bit-identical to being served alone.

Every ``segment`` steps the largest solute cluster (union-find over O-N contacts below
``rcontact``) is recorded for each request.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import system as S
from .host import PAIR_BLOCK, PREC_MIXED, Engine, ReplicaSpec


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the stream starts
    system: S.SofiaSystem
    seed: int
    temp: float
    nsteps: int = 500
    gamma_ps: float = 1.0


@dataclass
class Result:
    rid: int
    atoms: int
    arrival: float
    admitted: float
    finished: float
    clusters: list = field(default_factory=list)
    energies: np.ndarray | None = None
    state: tuple | None = None

    @property
    def latency_ms(self) -> float:
        return 1e3 * (self.finished - self.arrival)


@dataclass
class _Slot:
    req: Request
    x: np.ndarray
    v: np.ndarray
    step: int
    done: int
    admitted: float
    clusters: list
    f: np.ndarray | None = None      # total force after the last step (None before the first)


class SegmentServer:
    """Admission budget: hbm_budget_gb, or (None) 85 % of the GPU memory free when the server
    starts, read from the device.  A request's footprint is its atoms times bytes_per_atom;
    the first engine the server builds measures the real figure (device memory before and
    after) and replaces the initial estimate when it is larger."""

    def __init__(self, hbm_budget_gb: float | None = None, segment: int = 100,
                 precision: int = PREC_MIXED, bytes_per_atom: int = 6000,
                 rcontact: float = 3.5, keep_state: bool = False, tuning=None,
                 policy: str = "cohort", max_cohorts: int = 12):
        if hbm_budget_gb is None:
            from ._lib import device_memory
            self.budget = 0.85 * device_memory()[0]
        else:
            self.budget = hbm_budget_gb * 1e9
        self.segment = segment
        self.precision = precision
        self.bytes_per_atom = bytes_per_atom
        self._calibrated = False
        self.rcontact = rcontact
        self.keep_state = keep_state
        self.tuning = tuning
        self.segments_run = 0
        self.rebatches = 0
        if policy not in ("cohort", "rebatch"):
            raise ValueError("policy must be 'cohort' or 'rebatch'")
        self.policy = policy
        self.max_cohorts = max_cohorts
        self.cohorts_started = 0

    def footprint(self, req: Request) -> float:
        return req.system.np * self.bytes_per_atom

    def _engine(self, slots):
        reps = [ReplicaSpec(s.req.system, temp=s.req.temp, gamma_ps=s.req.gamma_ps, seed=s.req.seed)
                for s in slots]
        if not self._calibrated:
            from ._lib import device_memory
            before = device_memory()[0]
        eng = Engine(reps, precision=self.precision)
        if self.tuning:
            skin, lanes, waves = self.tuning[:3]
            eng.set_skin(skin)
            eng.set_tuning(lanes, PAIR_BLOCK, waves)
            if len(self.tuning) > 3:
                eng.set_bonded_mode(self.tuning[3])
            if len(self.tuning) > 4:
                eng.set_graph_steps(self.tuning[4])
        eng.put_state([(s.x, s.v) for s in slots])
        eng.set_steps([s.step for s in slots])
        eng.resume([s.f for s in slots])
        if not self._calibrated:
            used = before - device_memory()[0]
            natoms = sum(s.req.system.np for s in slots)
            if used > 0 and natoms > 0:
                self.bytes_per_atom = max(self.bytes_per_atom, used / natoms)
            self._calibrated = True
        return eng

    def serve(self, requests, realtime: bool = True):
        """Run the stream; returns (results, wall seconds, atom-steps).  With realtime=False
        the clock is virtual: one unit per segment."""
        if self.policy == "rebatch":
            return self._serve_rebatch(requests, realtime)
        return self._serve_cohorts(requests, realtime)

    def _result(self, s, now, energies, state):
        return Result(s.req.rid, s.req.system.np, s.req.arrival, s.admitted, now, s.clusters,
                      energies, state)

    def _serve_cohorts(self, requests, realtime: bool):
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        virtual = 0.0
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        cohorts = []                     # [engine, slots]
        results = []
        used = 0.0
        atom_steps = 0
        while pending or cohorts:
            now = clock() if realtime else virtual
            new = []
            while pending and pending[0].arrival <= now and used + self.footprint(pending[0]) <= self.budget:
                r = pending.popleft()
                new.append(_Slot(r, r.system.pos.astype(np.float64), r.system.v.astype(np.float64),
                                 0, 0, now, []))
                used += self.footprint(r)
            if new:
                cohorts.append([self._engine(new), new])
                self.cohorts_started += 1
            if len(cohorts) > self.max_cohorts:
                slots = [s for _, sl in cohorts for s in sl]
                for eng, sl in cohorts:
                    self._harvest(eng, sl)
                    eng.close()
                cohorts = [[self._engine(slots), slots]]
                self.rebatches += 1          # merged
            if not cohorts:
                if realtime:
                    time.sleep(max(0.0, pending[0].arrival - clock()))
                else:
                    virtual = pending[0].arrival
                continue
            steps = [min(self.segment, min(s.req.nsteps - s.done for s in sl)) for _, sl in cohorts]
            for (eng, _), n in zip(cohorts, steps):
                eng.launch(n)
            for eng, _ in cohorts:
                eng.finish()
            self.segments_run += 1
            for (eng, sl), n in zip(cohorts, steps):
                atom_steps += n * sum(s.req.system.np for s in sl)
                clusters = eng.largest_cluster(self.rcontact)
                for k, s in enumerate(sl):
                    s.done += n
                    s.clusters.append(int(clusters[k]))
            if not realtime:
                virtual += 1.0
            now = clock() if realtime else virtual
            keep = []
            for eng, sl in cohorts:
                fin = [k for k, s in enumerate(sl) if s.done >= s.req.nsteps]
                if not fin:
                    keep.append([eng, sl])
                    continue
                energies = eng.energies()
                states = eng.state() if self.keep_state else None
                for k in fin:
                    results.append(self._result(sl[k], now, energies[k], states[k] if states else None))
                    used -= self.footprint(sl[k].req)
                rest = [s for k, s in enumerate(sl) if k not in set(fin)]
                if rest:
                    self._harvest(eng, sl)
                    eng.close()
                    keep.append([self._engine(rest), rest])
                    self.rebatches += 1
                else:
                    eng.close()
            cohorts = keep
        wall = (clock() if realtime else virtual)
        return sorted(results, key=lambda r: r.rid), wall, atom_steps

    def _serve_rebatch(self, requests, realtime: bool):
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        virtual = 0.0
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        active: list[_Slot] = []
        in_engine: list[_Slot] = []      # slots held by the current engine, in its replica order
        results: list[Result] = []
        eng = None
        dirty = False
        used = 0.0
        atom_steps = 0
        while pending or active:
            now = clock() if realtime else virtual
            while pending and pending[0].arrival <= now and used + self.footprint(pending[0]) <= self.budget:
                r = pending.popleft()
                active.append(_Slot(r, r.system.pos.astype(np.float64), r.system.v.astype(np.float64),
                                    0, 0, now, []))
                used += self.footprint(r)
                dirty = True
            if not active:
                if realtime:
                    time.sleep(max(0.0, pending[0].arrival - clock()))
                else:
                    virtual = pending[0].arrival
                continue
            if dirty:
                if eng is not None:
                    self._harvest(eng, in_engine)
                    eng.close()
                eng = self._engine(active)
                self.rebatches += 1
                in_engine = list(active)
                dirty = False
            nsteps = min(self.segment, min(s.req.nsteps - s.done for s in active))
            eng.run(nsteps)
            self.segments_run += 1
            atom_steps += nsteps * sum(s.req.system.np for s in active)
            clusters = eng.largest_cluster(self.rcontact)
            for k, s in enumerate(active):
                s.done += nsteps
                s.clusters.append(int(clusters[k]))
            if not realtime:
                virtual += 1.0          # virtual clock: one unit per segment
            finished = [k for k, s in enumerate(active) if s.done >= s.req.nsteps]
            if finished:
                now = clock() if realtime else virtual
                energies = eng.energies()
                states = eng.state() if self.keep_state else None
                for k in finished:
                    s = active[k]
                    results.append(Result(s.req.rid, s.req.system.np, s.req.arrival, s.admitted, now,
                                          s.clusters, energies[k], states[k] if states else None))
                    used -= self.footprint(s.req)
                keep = [k for k in range(len(active)) if k not in set(finished)]
                if keep:
                    self._harvest(eng, in_engine)
                active = [active[k] for k in keep]
                eng.close()
                eng = None
                in_engine = []
                dirty = bool(active)
        if eng is not None:
            eng.close()
        wall = (clock() if realtime else virtual)
        return sorted(results, key=lambda r: r.rid), wall, atom_steps

    @staticmethod
    def _harvest(eng, slots):
        states = eng.state()
        forces = eng.forces()
        steps = eng.steps()
        for k, s in enumerate(slots):
            s.x, s.v = states[k]
            s.f = forces[k]
            s.step = int(steps[k])


def request_pool(sizes, seed: int = 20261206, minimize_steps: int = 150, precision=PREC_MIXED):
    """Pre-built, energy-minimised systems that requests draw their coordinates from."""
    rng = np.random.default_rng(seed)
    systems = []
    for n in sizes:
        x = rng.uniform(0.01, 0.04)
        g = int(round(x * n / (10 * x + 3 * (1 - x))))
        w = int(round((n - 10 * g) / 3))
        systems.append(S.make_system(g, w, seed=int(rng.integers(1 << 31))))
    relax(systems, minimize_steps, precision)
    return systems


def relax(systems, steps: int = 150, precision=PREC_MIXED):
    """Capped steepest descent of all systems in one batch; positions updated in place."""
    eng = Engine(systems, precision=precision)
    try:
        eng.init()
        eng.minimize(steps)
        for s, (x, _) in zip(systems, eng.state()):
            L = float(s.box)
            p = np.mod(x, L).astype(np.float32)
            p = np.where(p <= 0, p + np.float32(L), p)
            s.pos = p.astype(np.float32)
    finally:
        eng.close()
    return systems


def poisson_stream(pool, rate: float, duration: float, nsteps: int = 500, seed: int = 7,
                   t_lo: float = 280.0, t_hi: float = 320.0):
    """Requests with exponential inter-arrival times; each picks a pool system, a seed and a
    temperature, and gets Maxwell-Boltzmann velocities at that temperature."""
    rng = np.random.default_rng(seed)
    t, rid, out = 0.0, 0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        base = pool[int(rng.integers(len(pool)))]
        temp = float(rng.uniform(t_lo, t_hi))
        sd = int(rng.integers(1, 1 << 62))
        s = base.copy()
        s.v = S.maxwell_boltzmann(s, temp, np.random.default_rng(sd))
        out.append(Request(rid, t, s, sd, temp, nsteps))
        rid += 1
    return out
