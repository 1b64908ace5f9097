#!/usr/bin/env python
"""Throughput benchmark of the MI355X hot path (one JSON line on stdout).

Metric: aggregate atom-steps/s of batched replicas vs the plain one-system-per-launch
path, at a mixed-precision force error <= 1e-5 (relative RMS vs FP64), plus p50/p99
latency of served MD-segment requests.

* plain   -- the same 64 replicas advanced one system per launch (a single-replica engine
             each, same precision and tuning, resumed from the batched run's state):
             `batched_over_plain` is the gain of batching.
* plain_c2 -- 1,200 glycine + 64,000 water (204,000 atoms), FP64 pair arithmetic, FP64
             state, device-resident engine (one system).
* plain_dropin_c2 -- the same system through the drop-in C ABI, the reference's DSF step
             routine by routine on host arrays (PCIe copies per call); `plain_over_cpu`
             compares it with the CPU baseline on the same system and state.
* batched -- 64 replicas, 5,000-60,000 atoms (log-uniform, seeded), glycine mole
             fraction 0.01-0.04, FP32 pair arithmetic with int64 fixed-point sums, FP64
             state; all replicas advance in one launch sequence (hipGraph-captured step).
  `value` is the batched aggregate; a step = one velocity-Verlet step of every replica
  (list maintenance, LJ+DSF pair forces, 1-4 pairs, bonded terms, Langevin, SHAKE).
* served  -- Poisson arrivals of 500-step requests (3,000-150,000 atoms, 280-320 K),
             continuous batching, union-find cluster analysis every 100 steps.
* cpu_baseline -- the C oracle (OpenMP) on the 204,000-atom system at the plain run's
             equilibrated state, the same step, at least 20 steps; cpu_baseline_c1 the same
             on the 13,500-atom system.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Matom-steps/s batched vs plain @ force rel.err≤1e-5; p50/p99 segment ms"
FLOP_PER_PAIR = 42.0          # per ordered in-cutoff list entry (84 per unordered pair)
PEAK_FP32 = 157.3e12          # MI355X FP32 vector peak with packed FP32 (v_pk_fma_f32), FLOP/s
PEAK_FP32_UNPACKED = 78.6e12  # the same without packed FP32 (this build disables it)
PEAK_FP64 = 78.6e12
PAIR_KERNEL = "k_dsf_mixed pair kernel"


def _profile_json(name):
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def traffic_bytes():
    """HBM bytes per pair-kernel launch on the batched workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by the calibration copy), or
    None when no such summary exists."""
    d = _profile_json("pair_kernel_batched64_traffic.json")
    return None if d is None else d.get("traffic_bytes_per_launch")


def valu_counters():
    """VALU issue rate (instructions over the one-per-CU-per-cycle ceiling) and the FP32 /
    FP64 / integer instruction mix of the pair kernel on the batched workload, from the
    committed rocprofv3 counter summary, or None."""
    return _profile_json("pair_kernel_batched64_valu.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensure_built():
    from molecular_simulation_code_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from molecular_simulation_code_amd.build import build
        build()
    oracle_lib = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(oracle_lib):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def _make(args):
    from molecular_simulation_code_amd import system as S
    g, w, seed = args
    return S.make_system(g, w, seed=seed)


def make_systems(specs, workers):
    """Builds the systems (in a process pool when workers > 1), reporting progress on
    stderr every 8 systems so that a long generation is seen to advance."""
    out = []
    if workers > 1 and len(specs) > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(workers) as pool:
            for s in pool.imap(_make, specs):
                out.append(s)
                if len(out) % 8 == 0:
                    log(f"[bench] generated {len(out)}/{len(specs)} systems")
        return out
    for s in specs:
        out.append(_make(s))
        if len(out) % 8 == 0:
            log(f"[bench] generated {len(out)}/{len(specs)} systems")
    return out


def prepare_engine(eng, minimize, equil, tune=True):
    """Untimed set-up of a benchmark engine: capped steepest descent, Langevin
    equilibration at strong friction (gamma 20/ps) back to the replicas' temperatures, then
    the replicas' own friction and the autotuned skin / lanes.  Returns the tuning."""
    eng.init()
    eng.minimize(minimize)
    if equil > 0:
        eng.set_thermostat(gammas=[20.0] * len(eng.reps))
        eng.run(equil)
        eng.set_thermostat()
    return eng.autotune() if tune else None


def time_engine(eng, steps, warmup):
    eng.run(max(1, warmup))
    eng.synchronize()
    t0 = time.perf_counter()
    eng.run(steps)
    eng.synchronize()
    return time.perf_counter() - t0


def force_error_mixed_vs_fp64(system):
    """Relative RMS of the mixed-precision pair forces against the FP64 kernel."""
    from molecular_simulation_code_amd.host import SofiaGPU
    g = SofiaGPU(system)
    try:
        g.neighbor_wrapper(0)
        g.lj_cut_coul_dsf_nonewton(0)
        f64 = g.ff64.copy()
        g.lj_cut_coul_dsf_nonewton_mixed(0)
        fm = g.ff64.copy()
    finally:
        g.close()
    return float(np.sqrt(((fm - f64) ** 2).sum() / (f64 ** 2).sum()))


def cpu_model():
    """The host CPU's model name and logical CPU count, as /proc/cpuinfo reports them."""
    name, n = "unknown", 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and name == "unknown":
                    name = line.split(":", 1)[1].strip()
                if line.startswith("processor"):
                    n += 1
    except OSError:
        pass
    return name, n


def wrapped_f32(x, box):
    """FP64 positions wrapped into (0, box] as pbc() leaves them, in real*4."""
    p = np.mod(x, float(box)).astype(np.float32)
    p = np.where(p <= 0, p + np.float32(box), p).astype(np.float32)
    return np.where(p > np.float32(box), p - np.float32(box), p).astype(np.float32)


def dropin_rate(system, state, steps, warmup):
    """The plain GPU drop-in: the reference's DSF time loop (Source/mod_integrate.f90:270-308,
    force_wrapper_DSF_NOMIC Source/mod_force.f90:508-551) with every routine of the step
    replaced by its C-ABI drop-in on host-owned real*4 arrays, as a Fortran host would call
    them through the ISO_C_BINDING shim: nve_v, nve_x, pbc, check_neighbor (and
    neighbor_wrapper when triggered), force_wrapper_DSF_NOMIC (the FP64 DSF pair kernel, the
    1-4 pairs and the bonded terms), shake, nve_v.  Each routine copies the arrays it reads in and the arrays it writes
    out over PCIe (host.DropinStep, page-locked arrays), as the reference's MIC offload
    copied position in and ff out per call.  Returns (atom-steps/s, list builds during the
    timed steps, MB moved over the host link per step)."""
    from molecular_simulation_code_amd.host import DropinStep, SofiaGPU
    s = system.copy()
    s.pos = wrapped_f32(state[0], s.box)
    s.v = np.ascontiguousarray(state[1], np.float32)
    g = SofiaGPU(s)
    d = None
    try:
        g.neighbor_wrapper(0)
        d = DropinStep(g)
        d.forces()
        for _ in range(max(1, warmup)):
            d.step()
        builds = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            builds += d.step()
        el = time.perf_counter() - t0
        # per step without a rebuild, in units of np floats: nve_v (v, ff in; v out) twice,
        # nve_x, pbc, check, the force wrapper (position in, ff out) and shake
        floats = 2 * 9 + (4 + 3 + 3 + 4 + 3) + (4 + 4) + 3 + (4 + 3) + (4 + 3 + 3 + 3)
        return s.np * steps / el, builds, floats * 4 * s.np / 1e6
    finally:
        if d is not None:
            d.close()
        g.close()


def cpu_baseline(system, budget_s=12.0, min_steps=20, max_steps=1000):
    """Oracle MD steps (list rebuild when triggered, DSF pair forces, 1-4 pairs, bond, angle,
    torsion and improper terms, SHAKE, velocity Verlet) on the host cores: at least
    min_steps steps, then until budget_s has passed or max_steps are done.  Returns
    (atom-steps/s, steps, threads)."""
    from oracle import oracle as O
    from molecular_simulation_code_amd import system as S
    from molecular_simulation_code_amd.host import torsion_list
    s = system.copy()
    d = s.derived()
    prm = O.params_from(d, S.NTYPES)
    lj = S.lj_tables()
    mass = S.type_mass()
    kb = np.array([c[0] for c in S.BOND_COEFF]); r0 = np.array([c[1] for c in S.BOND_COEFF])
    ka = np.array([c[0] for c in S.ANGLE_COEFF]); th = np.radians([c[1] for c in S.ANGLE_COEFF])
    tk = np.array([c[0] for c in S.DIHED_COEFF]); tm = np.array([c[1] for c in S.DIHED_COEFF], float)
    ts = np.array([c[2] for c in S.DIHED_COEFF], float)
    ik = np.array([c[0] for c in S.IMPRO_K_CHI0]); ic = np.array([c[1] for c in S.IMPRO_K_CHI0])
    rn, nT, nD = O.cell_grid(d["box"], d["rcut"], d["cut_coul"])
    cnum = O.neighcell(nT, nD)
    st = {"tors": torsion_list(s).copy(), "impro": s.improlist.copy()}

    def rebuild():
        tt, perm, start, endp, atombin = O.sort_cells(s.pos, s.types, d["box"], rn, nT, nD)
        inv = perm - 1
        s.pos, s.types, s.q, s.v = s.pos[inv], s.types[inv], s.q[inv], s.v[inv]
        sp = s.specbond[inv]
        s.specbond = np.where(sp > 0, tt[np.maximum(sp, 1) - 1], 0).astype(np.int32)
        s.num3bond = s.num3bond[inv]
        s.bondlist[:, :2] = tt[s.bondlist[:, :2] - 1]
        s.anglelist[:, :3] = tt[s.anglelist[:, :3] - 1]
        s.dihedlist[:, :4] = tt[s.dihedlist[:, :4] - 1]
        st["tors"][:, :4] = tt[st["tors"][:, :4] - 1]
        st["impro"][:, :4] = tt[st["impro"][:, :4] - 1]
        sa = s.shake_atoms
        s.shake_atoms = np.where(sa > 0, tt[np.maximum(sa, 1) - 1], 0).astype(np.int32)
        st["nl"], st["nn"], st["na"] = O.build_neighbor(s.pos, s.types, atombin, start, endp, cnum,
                                                        s.num3bond, s.specbond, d["box"], d["ibox"],
                                                        d["rv2"], st.get("na", S.NEIGH_ALLOC0))
        st["dr"] = [np.zeros(s.np, np.float32) for _ in range(3)]

    def forces():
        ff, _, _, _, _ = O.dsf(s.pos, s.types, s.q, st["nl"], st["nn"], st["na"], lj, prm)
        fh, _, _ = O.pair14_dsf(s.dihedlist, s.pos, s.types, s.q, lj, prm)
        fb, _ = O.bond_harmonic(s.bondlist, s.pos, s.types, kb, r0, d["box"], d["ibox"])
        fa, _ = O.angle_harmonic(s.anglelist, s.pos, s.types, ka, th, d["box"], d["ibox"])
        ft, _ = O.torsion_cos(st["tors"], s.pos, s.types, tk, tm, ts, d["box"], d["ibox"])
        fi, _, _ = O.improper_harmonic(st["impro"], s.pos, s.types, ik, ic, d["box"], d["ibox"])
        f = (ff.astype(np.float64) + fh + fb + fa + ft + fi).astype(np.float32)
        f, _, _, _ = O.shake(1, s.pos, s.types, s.v, f, mass, d["dt"], S.KCAL2AMU, d["box"],
                             d["ibox"], s.shake_num, s.shake_atoms, s.shake_bond,
                             float(S.SHAKE_TOL), S.SHAKE_MAXITER)
        return f

    def step(f):
        s.v, _ = O.nve_v(d["dthalf"], s.pos, s.types, mass, s.v, f, float(S.AMU2E))
        s.pos, *st["dr"] = O.nve_x(d["dt"], s.v, s.pos, s.types, *st["dr"])
        s.pos = O.pbc(s.pos, s.types, d["box"])
        if O.check_neighbor(*st["dr"], d["rvrc2"]):
            rebuild()
        f = forces()
        s.v, _ = O.nve_v(d["dthalf"], s.pos, s.types, mass, s.v, f, float(S.AMU2E))
        return f

    rebuild()
    f = forces()
    f = step(f)                                   # warm-up step
    t0 = time.perf_counter()
    n = 0
    while True:
        f = step(f)
        n += 1
        el = time.perf_counter() - t0
        if n >= max_steps or (n >= min_steps and el > budget_s):
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return s.np * n / el, n, threads


def drift_run(args, workers):
    """--drift-config c3-mixed: 6,000 glycine + 320,000 water (1,020,000 atoms), mixed
    precision; c2-fp64: 1,200 glycine + 64,000 water (204,000 atoms), FP64.  SHAKE, dt 2 fs:
    minimise, Langevin-equilibrate, then --drift-steps NVE steps with the total energy
    sampled every 250 steps, then --drift-steps NVT steps (Langevin 1/ps).  Drift is the
    least-squares slope of the total energy, per ns and per atom."""
    from molecular_simulation_code_amd import system as S
    from molecular_simulation_code_amd import host
    t = time.perf_counter()
    if args.drift_config == "c2-fp64":
        spec, prec, what = (1200, 64000, 20261203), host.PREC_FP64, "1,200 Gly + 64,000 water), FP64"
    else:
        spec, prec, what = (6000, 320000, 20261204), host.PREC_MIXED, "6,000 Gly + 320,000 water), mixed precision"
    s = make_systems([spec], 1)[0]
    log(f"[drift] {s.np} atoms, box {float(s.box):.2f} A, generated in {time.perf_counter() - t:.1f}s")
    eng = host.Engine([host.ReplicaSpec(s, temp=300.0, gamma_ps=20.0, seed=11)], precision=prec)
    out = {"workload": f"{s.np}-atom glycine/water ({what}, SHAKE, dt 2 fs", "atoms": s.np}
    try:
        eng.init()
        eng.minimize(args.minimize)
        eng.run(max(0, args.equil))                      # Langevin at 20/ps
        out["tuning"] = list(eng.autotune(steps=40))
        eng.set_thermostat(gammas=[0.0])                 # NVE
        sample, series = 250, []
        eng.run(1)
        e = eng.energies()[0]
        series.append((0, float(e.sum()), float(e[7])))
        t0 = time.perf_counter()
        for k in range(args.drift_steps // sample):
            eng.run(sample)
            e = eng.energies()[0]
            series.append(((k + 1) * sample, float(e.sum()), float(e[7])))
            log(f"[drift] NVE step {(k + 1) * sample}: E {e.sum():.3f} kcal/mol, KE {e[7]:.3f}")
        t_nve = time.perf_counter() - t0
        st = np.array([x[0] for x in series], float)
        et = np.array([x[1] for x in series], float)
        ke = np.array([x[2] for x in series], float)
        dt_ns = float(s.derived()["dt"]) * 1e-6
        slope = np.polyfit(st * dt_ns, et, 1)[0]         # kcal/mol per ns
        out["nve"] = {"steps": int(st[-1]), "sample_every": sample,
                      "E_start": et[0], "E_end": et[-1], "KE_mean": float(ke.mean()),
                      "drift_kcal_per_mol_per_ns": slope,
                      "drift_kcal_per_mol_per_ns_per_atom": slope / s.np,
                      "drift_per_ns_relative_to_KE": slope / float(ke.mean()),
                      "E_rms_fluctuation": float(np.std(et - np.polyval(np.polyfit(st, et, 1), st))),
                      "temperature_K_end": float(eng.temperatures()[0]),
                      "ms_per_step": 1e3 * t_nve / max(1, int(st[-1]))}
        eng.set_thermostat(temps=[300.0], gammas=[1.0])  # NVT
        t0 = time.perf_counter()
        eng.run(args.drift_steps)
        out["nvt"] = {"steps": args.drift_steps, "temperature_K_end": float(eng.temperatures()[0]),
                      "ms_per_step": 1e3 * (time.perf_counter() - t0) / args.drift_steps}
        out["series"] = series
    finally:
        eng.close()
    print(json.dumps(out))


def profile_run(args, workers):
    """The batched workload for a profiler, prepared as the timed run prepares it
    (minimisation, equilibration, autotuning), then a few MD steps, 5 launches of the pair
    kernel and 5 of the 1 GiB streaming copy whose traffic is known (2 GiB per launch)."""
    from molecular_simulation_code_amd import system as S
    import pickle
    if args.cache and os.path.exists(args.cache):
        with open(args.cache, "rb") as f:      # written by this script
            reps = pickle.load(f)
    else:
        comps = S.replica_sizes(args.replicas)
        reps = make_systems([(g, w, 20261210 + k) for k, (g, w) in enumerate(comps)], workers)
        if args.cache:
            with open(args.cache, "wb") as f:
                pickle.dump(reps, f)
    if args.generate_only:
        return
    from molecular_simulation_code_amd import host
    rng = np.random.default_rng(5)
    specs = [host.ReplicaSpec(r, temp=float(rng.uniform(280, 320)), gamma_ps=1.0, seed=100 + k)
             for k, r in enumerate(reps)]
    eng = host.Engine(specs, precision=host.PREC_MIXED)
    try:
        if args.tuning:   # the timed run's tuning, given instead of autotuned
            knobs = [float(x) for x in args.tuning.split(",")]
            eng.set_skin(knobs[0])
            eng.set_tuning(int(knobs[1]), host.PAIR_BLOCK, int(knobs[2]))
            if len(knobs) > 3:
                eng.set_bonded_mode(int(knobs[3]))
            if len(knobs) > 4:
                eng.set_graph_steps(int(knobs[4]))
            # the timed run's preparation, with the equilibration stepped eagerly so that the
            # profiled process captures one graph only
            eng.init()
            eng.minimize(args.minimize)
            log(f"[profile-run] minimised ({args.minimize} steps)")
            if args.equil > 0:
                eng.set_thermostat(gammas=[20.0] * len(eng.reps))
                for k in range(0, args.equil, 100):
                    eng.run(min(100, args.equil - k), use_graph=False)
                    log(f"[profile-run] equilibration {min(k + 100, args.equil)}/{args.equil}")
                eng.set_thermostat()
            tuning = args.tuning
        else:
            tuning = prepare_engine(eng, args.minimize, args.equil, tune=not args.no_tune)
        eng.run(max(1, args.steps))
        ms, pairs = eng.profile_nonbond(5)
        ms_copy = eng.profile_copy(1 << 30, 5)
        log(f"[profile-run] {sum(r.np for r in reps)} atoms, tuning {tuning}; pair kernel "
            f"{ms:.4f} ms/launch, {pairs} entries in cutoff; copy 1 GiB {ms_copy:.4f} ms/launch")
        # the last 5 list-build launches of the run rebuild every replica
        full = eng.profile_step(5, full_rebuild=True)
        log(f"[profile-run] full rebuild of every replica: list build {full['list build']:.4f} ms, "
            f"sort+gather {full['sort+gather']:.4f} ms")
    finally:
        eng.close()


def dropin_run(args):
    """C2 (204,000 atoms) minimised and equilibrated on the FP64 engine as the full run does,
    then args.steps drop-in steps after args.warmup; logs the rate."""
    from molecular_simulation_code_amd import _lib, host
    from molecular_simulation_code_amd import system as S
    c2 = S.make_system(1200, 64000, seed=20261203)
    if _lib.device_count() < 1:
        log("no HIP device")
        sys.exit(3)
    eng = host.Engine([host.ReplicaSpec(c2, temp=300.0, gamma_ps=1.0, seed=1)],
                      precision=host.PREC_FP64)
    eng.init()                       # prepare_engine's steps, eager (no graph under a tracer)
    eng.minimize(args.minimize)
    eng.set_thermostat(gammas=[20.0])
    eng.run(args.equil, use_graph=False)
    state = eng.state()[0]
    eng.close()
    log(f"[plain drop-in] C2 prepared: minimise {args.minimize}, Langevin {args.equil} steps")
    rate, builds, mb = dropin_rate(c2, state, args.steps, args.warmup)
    log(f"[plain drop-in] {rate / 1e6:.1f} Matom-steps/s ({args.steps} steps, {builds} list "
        f"builds, {mb:.0f} MB over the host link per step)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--minimize", type=int, default=150)
    ap.add_argument("--equil", type=int, default=400,
                    help="untimed Langevin steps at gamma 20/ps after the minimisation")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--cache", default=None,
                    help="profile run: pickle of the generated replicas (written if absent)")
    ap.add_argument("--generate-only", action="store_true",
                    help="profile run: only generate the replicas into --cache")
    ap.add_argument("--drift", action="store_true",
                    help="energy-conservation run on the 1.02 M-atom mixed system: NVE segment of "
                         "--drift-steps steps, then as many NVT steps; prints one JSON line")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--drift-config", choices=["c3-mixed", "c2-fp64"], default="c3-mixed",
                    help="drift run: the 1.02 M-atom mixed system or the 204,000-atom FP64 one")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run of the batched workload for rocprofv3: a few steps, the pair "
                         "kernel and the 1 GiB calibration copy; no JSON line")
    ap.add_argument("--dropin-run", action="store_true",
                    help="the plain drop-in leg alone for rocprofv3: C2 prepared as the full run "
                         "prepares it, then --steps drop-in steps; no JSON line")
    ap.add_argument("--tuning", default=None,
                    help="profile run: skin,lanes,list_waves[,bonded_mode[,steps_per_graph]] to use "
                         "instead of the autotuner")
    ap.add_argument("--workers", type=int, default=0,
                    help="input-generation processes (0 = one per core up to 16; use 1 under a "
                         "profiler, which initialises the GPU before the program starts)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); replicas shard "
              "across GPUs with no communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    ensure_built()
    from molecular_simulation_code_amd import system as S
    workers = args.workers if args.workers > 0 else min(16, os.cpu_count() or 1)
    K, W = args.steps, args.warmup

    if args.profile_run:
        profile_run(args, workers)
        return
    if args.drift:
        drift_run(args, workers)
        return
    if args.dropin_run:
        dropin_run(args)
        return

    # ---------------------------------------------------------------- inputs (CPU, before any GPU use)
    t = time.perf_counter()
    comps = S.replica_sizes(args.replicas)
    specs_cpu = [(1200, 64000, 20261203)] + [(g, w, 20261210 + k) for k, (g, w) in enumerate(comps)]
    pool_sizes = []
    if not args.no_serve:
        lo, hi = 3000, 150000
        pool_sizes = [int(math.exp(x)) for x in np.linspace(math.log(lo), math.log(hi), 10)]
        prng = np.random.default_rng(20261206)
        for n in pool_sizes:
            x = prng.uniform(0.01, 0.04)
            g = int(round(x * n / (10 * x + 3 * (1 - x))))
            specs_cpu.append((g, int(round((n - 10 * g) / 3)), int(prng.integers(1 << 31))))
    specs_cpu.append((150, 4000, 20261202))
    log(f"[inputs] generating {len(specs_cpu)} systems with {workers} workers")
    systems = make_systems(specs_cpu, workers)
    plain_sys, reps = systems[0], systems[1:1 + len(comps)]
    pool = systems[1 + len(comps):1 + len(comps) + len(pool_sizes)]
    small = systems[-1]
    log(f"  done in {time.perf_counter() - t:.1f}s")

    from molecular_simulation_code_amd import _lib, host
    from molecular_simulation_code_amd import server as SV
    if _lib.device_count() < 1:
        log("no HIP device")
        sys.exit(3)

    # ---------------------------------------------------------------- plain: 204k FP64
    log(f"[plain] {plain_sys.np} atoms, FP64")
    eng = host.Engine([host.ReplicaSpec(plain_sys, temp=300.0, gamma_ps=1.0, seed=1)],
                      precision=host.PREC_FP64)
    tune_p = prepare_engine(eng, args.minimize, args.equil, not args.no_tune)
    log(f"[plain] tuned (skin, lanes, ms/step) = {tune_p}, T = {eng.temperatures()[0]:.1f} K")
    dt_plain = time_engine(eng, K, W)
    plain_rate = plain_sys.np * K / dt_plain
    ms_p, pairs_p = eng.profile_nonbond(10)
    temp_p = float(eng.temperatures()[0])
    phases_p = eng.profile_step(30)
    state_c2 = eng.state()[0]
    log(f"[plain] {plain_rate / 1e6:.1f} Matom-steps/s  ({1e3 * dt_plain / K:.3f} ms/step, "
        f"pair kernel {ms_p:.3f} ms, {pairs_p / plain_sys.np:.1f} pairs/atom in cutoff, T {temp_p:.1f} K)")
    eng.close()

    # ---------------------------------------------------------------- plain drop-in: C2 through the C ABI
    kd = max(10, min(K, 100))
    dropin, dropin_builds, dropin_mb = dropin_rate(plain_sys, state_c2, kd, max(1, W // 2))
    log(f"[plain drop-in] {dropin / 1e6:.1f} Matom-steps/s ({kd} steps, {dropin_builds} list builds, "
        f"host arrays, {dropin_mb:.0f} MB over the host link per step)")

    # ---------------------------------------------------------------- batched: 64 replicas mixed
    natoms = sum(r.np for r in reps)
    log(f"[batched] {len(reps)} replicas, {natoms} atoms, mixed precision")
    rng = np.random.default_rng(5)
    specs = [host.ReplicaSpec(r, temp=float(rng.uniform(280, 320)), gamma_ps=1.0, seed=100 + k)
             for k, r in enumerate(reps)]
    eng = host.Engine(specs, precision=host.PREC_MIXED)
    tune_b = prepare_engine(eng, args.minimize, args.equil, not args.no_tune)
    if tune_b is not None:
        log(f"[batched] autotune log: {eng.tune_log}")
    log(f"[batched] tuned (skin, lanes, ms/step) = {tune_b}, "
        f"T = {eng.temperatures().mean():.1f} K (mean)")
    b0 = eng.rebuilds()
    dt_b = time_engine(eng, K, W)
    batched_rate = natoms * K / dt_b
    rebuilds = eng.rebuilds() - b0
    temps_b = eng.temperatures()
    ms_b, pairs_b = eng.profile_nonbond(10)
    ms_copy = eng.profile_copy(1 << 30, 10)
    snap = (eng.state(), eng.forces(), eng.steps())
    phases_b = eng.profile_step(30)
    log("[batched] phases (ms/step, eager): " + ", ".join(f"{k} {v:.3f}" for k, v in phases_b.items()))
    log(f"[batched] {batched_rate / 1e6:.1f} Matom-steps/s ({1e3 * dt_b / K:.3f} ms/step, "
        f"pair kernel {ms_b:.3f} ms, rebuilds/replica {rebuilds.mean():.1f} over {K + max(1, W)} steps, "
        f"T {temps_b.mean():.1f} K)")
    eng.close()

    # ---------------------------------------------------------------- plain: the same replicas, one per launch
    log(f"[plain/replica] the {len(reps)} replicas one system per launch, mixed precision")
    kp, wp = max(1, K // 2), max(1, W // 2)
    t_rep, as_rep = 0.0, 0
    for k, spec in enumerate(specs):
        e1 = host.Engine([spec], precision=host.PREC_MIXED)
        try:
            e1.put_state([snap[0][k]])
            e1.set_steps([snap[2][k]])
            if tune_b is not None:
                e1.set_skin(tune_b[0])
                e1.set_tuning(tune_b[1], host.PAIR_BLOCK, tune_b[3])
                e1.set_bonded_mode(tune_b[4])
                e1.set_graph_steps(tune_b[5])
            e1.resume([snap[1][k]])
            t_rep += time_engine(e1, kp, wp)
            as_rep += spec.system.np * kp
        finally:
            e1.close()
    per_launch_rate = as_rep / t_rep
    log(f"[plain/replica] {per_launch_rate / 1e6:.1f} Matom-steps/s ({kp} steps per replica)")

    # force error of the mixed kernel against FP64 on a replica-sized system
    ferr = force_error_mixed_vs_fp64(reps[0])
    log(f"[precision] mixed vs FP64 force rel RMS = {ferr:.2e}")

    # ---------------------------------------------------------------- served stream
    p50 = p99 = served_rate = serve_budget = None
    nreq = 0
    if not args.no_serve:
        log(f"[serve] relaxing a pool of {len(pool)} systems")
        SV.relax(pool, args.minimize)
        reqs = SV.poisson_stream(pool, args.serve_rate, args.serve_seconds)
        nreq = len(reqs)
        tuning = None if tune_b is None else (tune_b[0], tune_b[1], tune_b[3], tune_b[4], tune_b[5])
        srv = SV.SegmentServer(segment=100, tuning=tuning)
        res, wall, atom_steps = srv.serve(reqs)
        lat = np.array([r.latency_ms for r in res])
        p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
        served_rate = atom_steps / wall
        serve_budget = (srv.budget / 1e9, srv.bytes_per_atom)
        log(f"[serve] {nreq} requests, p50 {p50:.0f} ms, p99 {p99:.0f} ms, "
            f"{served_rate / 1e6:.1f} Matom-steps/s, {srv.cohorts_started} cohorts, "
            f"{srv.rebatches} re-batches, {srv.segments_run} segments; HBM budget "
            f"{serve_budget[0]:.0f} GB read from the device, {serve_budget[1]:.0f} B/atom measured")

    # ---------------------------------------------------------------- CPU baseline
    cpu = cpu_c1 = None
    if not args.no_cpu:
        model, ncpu = cpu_model()
        c2 = plain_sys.copy()
        c2.pos = wrapped_f32(state_c2[0], c2.box)
        c2.v = np.ascontiguousarray(state_c2[1], np.float32)
        rate, nsteps, threads = cpu_baseline(c2, args.cpu_seconds, min_steps=20)
        cpu = {"value": rate / 1e6, "unit": "Matom-steps/s", "cores": threads, "kind": "port",
               "cpu_model": model, "host_logical_cpus": ncpu,
               "sample": f"C oracle (OpenMP pair and list loops; DSF, 1-4, bond, angle, torsion, "
                         f"improper, SHAKE, velocity Verlet, list rebuilds when triggered) on the "
                         f"204,000-atom system at the plain run's equilibrated state, {nsteps} MD "
                         "steps"}
        log(f"[cpu] C2: {rate / 1e6:.3f} Matom-steps/s on {threads} threads ({model})")
        r1, n1, _ = cpu_baseline(small, args.cpu_seconds, min_steps=20, max_steps=1000)
        cpu_c1 = {"value": r1 / 1e6, "unit": "Matom-steps/s", "steps": n1,
                  "sample": "13,500-atom glycine/water (generated state), same step"}
        log(f"[cpu] C1: {r1 / 1e6:.3f} Matom-steps/s ({n1} steps)")

    flops = FLOP_PER_PAIR * pairs_b
    achieved = flops / (ms_b * 1e-3)
    valu = valu_counters()
    line = {
        "metric": METRIC,
        "value": round(batched_rate / 1e6, 3),
        "unit": "Matom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * dt_b / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {len(reps)} glycine/water replicas, 5,000-60,000 atoms "
                        f"log-uniform ({natoms} atoms), FP32 pair math + int64 fixed-point "
                        "sums, FP64 state, Langevin 1/ps, SHAKE, dt 2 fs, hipGraph steps",
            "replicas": len(reps), "atoms": natoms,
            "temperature_K_mean": round(float(temps_b.mean()), 1),
            "phase_ms_eager": {k: round(v, 4) for k, v in phases_b.items()},
            "tuning": None if tune_b is None else {"skin": tune_b[0], "lanes": tune_b[1],
                                                   "list_waves": tune_b[3], "bonded_mode": tune_b[4],
                                                   "steps_per_graph": tune_b[5]},
            "preparation": f"minimize {args.minimize}, Langevin gamma 20/ps for {args.equil} steps "
                           "(untimed)",
            "plain": {"workload": f"the same {len(reps)} replicas, one system per launch (one engine "
                                  f"each, hipGraph steps, mixed precision, batched tuning), {kp} "
                                  "timed steps each",
                      "value": round(per_launch_rate / 1e6, 3), "unit": "Matom-steps/s"},
            "plain_c2": {"workload": "204,000-atom glycine/water, FP64 pair math and state",
                      "value": round(plain_rate / 1e6, 3), "unit": "Matom-steps/s",
                      "ms_per_step": round(1e3 * dt_plain / K, 4), "dtype": "f64",
                      "temperature_K": round(temp_p, 1),
                      "phase_ms_eager": {k: round(v, 4) for k, v in phases_p.items()},
                      "tuning": None if tune_p is None else {"skin": tune_p[0], "lanes": tune_p[1],
                                                             "list_waves": tune_p[3],
                                                             "bonded_mode": tune_p[4],
                                                             "steps_per_graph": tune_p[5]}},
            "served": {"workload": f"Poisson {args.serve_rate:g}/s for {args.serve_seconds:g}s, "
                                   "500-step requests, 3,000-150,000 atoms",
                       "requests": nreq, "p50_ms": p50, "p99_ms": p99,
                       "hbm_budget_GB": None if serve_budget is None else round(serve_budget[0], 1),
                       "bytes_per_atom": None if serve_budget is None else round(serve_budget[1]),
                       "value": None if served_rate is None else round(served_rate / 1e6, 3)},
        },
        "batched_over_plain": round(batched_rate / per_launch_rate, 3),
        "batched_over_plain_c2_fp64": round(batched_rate / plain_rate, 3),
        # the plain GPU drop-in (C2 through the C ABI, host arrays) over the CPU port on C2
        "plain_dropin_c2": {"value": round(dropin / 1e6, 3), "unit": "Matom-steps/s",
                            "steps": kd, "list_builds": int(dropin_builds),
                            "host_link_MB_per_step": round(dropin_mb, 1),
                            "workload": "204,000-atom glycine/water, FP64: the reference's DSF step "
                                        "with each routine replaced by its C-ABI drop-in on host "
                                        "real*4 arrays in page-locked memory, each routine copying "
                                        "in what it reads and out what it writes (PCIe included)"},
        "plain_over_cpu": None if cpu is None else round(dropin / 1e6 / cpu["value"], 1),
        "plain_engine_c2_over_cpu": None if cpu is None else round(plain_rate / 1e6 / cpu["value"], 1),
        "cpu_baseline_c1": cpu_c1,
        "p50_ms": p50,
        "p99_ms": p99,
        "force_rel_rms_mixed_vs_fp64": ferr,
        "list_rebuilds_per_replica_mean": float(np.mean(rebuilds)),
        "list_rebuild_window_steps": K + max(1, W),
        "hbm_copy_GBps": round(2.0 * (1 << 30) / (ms_copy * 1e-3) / 1e9, 1),
        # the pair kernel is FP32 vector-ALU work (no matrix cores): priced against the
        # packed FP32 vector peak (frac) and the non-packed one this build issues (frac_unpacked)
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "peak_unpacked": PEAK_FP32_UNPACKED / 1e12,
                     "frac_unpacked": achieved / PEAK_FP32_UNPACKED, "traffic": traffic_bytes(),
                     "traffic_workload": (_profile_json("pair_kernel_batched64_traffic.json")
                                          or {}).get("workload"),
                     "valu_counters": valu,
                     "kernel": PAIR_KERNEL, "ms_per_launch": ms_b,
                     "pairs_in_cutoff_per_launch": pairs_b, "flop_per_pair": FLOP_PER_PAIR},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
