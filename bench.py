"""Benchmark: particle-steps/s of the per-step pipeline on one MI355X.

Default workload (one GPU): the 8M-particle skewed cell -- a periodic 2000 nm
box with a 10 nm cutoff (= cell), a lipid shell of radius 900 +- 5 nm
(2,097,152 lipids), a dense protein ball of radius 150 nm (1,258,291), and
uniform metabolite carriers and dummies (4,404,019 and 629,146); sub-iteration
counts C/R/D = 16/32/8.  value = (active plain + active metabolic particles) *
steps / seconds, with state resident in HBM and every step replayed from the
captured hipGraph, after the autotuner has picked the scheduling options.

Also reported: per-stage time and algorithmic-byte bandwidth, the roofline of
the dominant kernel pair (the relax pass) from HIP events on the library's
stream, the CPU oracle timed on a sub-box of the same structure (built for the
host it runs on), and the 256-cell batch against the same cells one at a time.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spatial-cell_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import scenes  # noqa: E402
import spatial_cell as sc  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TLANES = 256 * 4 * 16 * 2.4e9 / 1e12   # VALU lane-instructions per second, in T (39.3)


def stage_bytes(n_p, n_l, n_m, passes, P):
    """Algorithmic bytes per step for each stage: the fp32/int32 fields each
    stage must read and write, each field once per kernel, neighbour data once
    per pass; cell-table reads are not counted."""
    C, R, D = P.noise_coordination_steps, P.relaxation_steps, P.metabolite_diffusion_steps
    even = (R + 1) // 2
    return {
        # key (pos, flags -> key, index), histogram + passes over (key, index), permute of the
        # persistent state, cell range writes
        "sort": (24 + 4 + 16 * passes + 180 + 12) * n_p + (24 + 4 + 16 * passes + 228 + 12) * n_m,
        "move": 48 * (n_p + n_m),                    # Philox noise in, posNoise/angNoise/debug out
        "coordinate": 8 * n_p * C,                   # unbonded early-out (no bonds in this workload)
        "apply_noise": 84 * (n_p + n_m),
        "relax": 108 * n_p * R,                      # 80 B read + 28 B written per particle
        "relax_metabolic": (44 * n_m + 36 * n_l) * even,
        "relax_partners": (8 * n_p + 40 * n_l + 32 * n_m) * even,
        "apply_velocities": 112 * (n_p + n_m),
        "complexify": 52 * n_p,
        "transition": 16 * n_p,
        "polymerize": 0,
        "diffuse": 96 * n_m * D,
    }


def radix_passes(P):
    """8-bit LSD passes over the cell key plus the inactive-sentinel bit
    (grid of config.cuh:66-74)."""
    import math
    n = math.ceil(P.sim_size / P.grid_cell_size)
    bits, p2 = 1, 2
    while p2 < n:
        p2 *= 2
        bits += 1
    return (3 * bits + 1 + 7) // 8


def cpu_sample_scene(workload):
    """A sub-box of the workload with its structure and densities, sized for
    about a second of oracle time per step."""
    if workload == "c3":
        return scenes.c3_like_scene(0.08, init=False), "C3 structure scaled to a 160 nm periodic box"
    if workload == "c1":
        return scenes.c1_scene(init=False), "the C1 cell itself"
    return scenes.uniform_scene(16384, 250.0, boundary="periodic", init=False), \
        "a 250 nm periodic sub-box at the C2 density"


def native_oracles():
    """The oracle built with -O3 -march=native for this host (serial and
    OpenMP), once per run, into a temporary directory."""
    import importlib.util
    import tempfile
    spec = importlib.util.spec_from_file_location("orc_build", os.path.join(ROOT, "oracle", "build.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.build_native(tempfile.mkdtemp(prefix="sc_oracle_"))


def cpu_baseline(workload, lib_path, seconds_budget=12.0, threaded=False):
    """The oracle on a sub-box of the workload: init, then steps for about
    seconds_budget (at least 2), timed.  Single-threaded, or (threaded) the
    OpenMP build with its pure Jacobi passes over all host threads.  The
    first step is also run by the parity build of the oracle, and the two
    states must agree bit for bit (the host-tuned build computes the same)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    (cfg, setup), what = cpu_sample_scene(workload)
    s = sc.Setup(cfg, setup)
    init = ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1))
    o = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic, lib_path=lib_path)
    for st, arg in init:
        o.run_stage(st, arg)
    c = o.snapshot()[0]
    steps = 0
    t0 = time.perf_counter()
    while steps < 2 or (time.perf_counter() - t0 < seconds_budget and steps < 400):
        o.step(1)
        steps += 1
        if steps == 1:
            t_first = time.perf_counter()
            first = o.snapshot()
            t0 += time.perf_counter() - t_first          # the snapshot is not timed
    dt = time.perf_counter() - t0
    o.close()
    ref = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic)
    for st, arg in init:
        ref.run_stage(st, arg)
    ref.step(1)
    want = ref.snapshot()
    ref.close()
    same = first[0].as_tuple() == want[0].as_tuple() and first[1].tobytes() == want[1].tobytes() and \
        first[2].tobytes() == want[2].tobytes()
    n = c.n_active + c.n_active_metabolic
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if threaded else 1
    how = f"OpenMP build, {threads} threads" if threaded else "serial C++, one thread"
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "build": "g++ -O3 -march=native -ffp-contract=off", "same_bits_as_parity_build": bool(same),
            "sample": f"oracle ({how} of a {model}, {os.cpu_count()} logical CPUs) on {what} "
                      f"({n} particles), {steps} steps, {dt:.1f} s"}


def copy_probe(mib):
    """A plain device copy of known size: FETCH_SIZE / WRITE_SIZE of its
    kernel against mib MiB calibrate the counters for the roofline traffic."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}))


def batched_vs_plain(args, steps=None, warmup=None, emit=True):
    """256 independent C1-mix cells (sizes: tests/golden/c4_sizes.json, 2^24
    particles in all), stepped as one batch (one launch chain) and as 256
    single cells one after another (one launch chain each)."""
    sizes = scenes.c4_sizes()
    docs = [scenes.c1_mix_scene(n, seed=42 + i) for i, n in enumerate(sizes)]
    steps = max(1, args.steps if steps is None else steps)
    warmup = max(1, args.warmup if warmup is None else warmup)
    # the one-by-one form can be timed over fewer steps (it is ~8x slower):
    # both figures are particle-steps per second over their own timed steps
    plain_steps = max(1, min(steps, getattr(args, "plain_steps", 0) or steps))
    t = time.perf_counter()
    batch = sc.Batch([c for c, _ in docs], [s for _, s in docs])
    batch.synchronize()
    print(f"[c4] batch of {len(docs)} created in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    batch.step(warmup)
    batch.synchronize()
    n_particles = 0
    for i in range(len(docs)):
        c = batch.instance_counts(i)
        n_particles += c.n_active + c.n_active_metabolic
    # the first plain_steps steps timed on their own (the window the
    # one-by-one form is timed over: bonds accumulate, so later steps cost
    # more), then the rest
    t = time.perf_counter()
    batch.step(plain_steps)
    batch.synchronize()
    wall_head = time.perf_counter() - t
    if steps > plain_steps:
        batch.step(steps - plain_steps)
        batch.synchronize()
    wall_b = time.perf_counter() - t
    print(f"[c4] batched: {steps} steps in {wall_b:.3f}s (the first {plain_steps} in {wall_head:.3f}s)",
          file=sys.stderr, flush=True)
    small = int(np.argmin(sizes))
    snap_b = batch.instance_snapshot(small)
    batch.close()

    t = time.perf_counter()
    sims = [sc.Simulation(c, s) for c, s in docs]
    for sim in sims:
        sim.step(warmup)
    for sim in sims:
        sim.synchronize()
    print(f"[c4] {len(sims)} single cells created and warmed in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    t = time.perf_counter()
    for sim in sims:                          # one launch chain per cell, one cell after another
        sim.step(plain_steps)
        sim.synchronize()
    wall_p = time.perf_counter() - t
    print(f"[c4] one by one: {plain_steps} steps in {wall_p:.3f}s", file=sys.stderr, flush=True)
    if plain_steps < steps:                   # the same step count before comparing the smallest cell
        sims[small].step(steps - plain_steps)
    snap_p = sims[small].snapshot()
    same = all(a.tobytes() == b.tobytes() for a, b in zip(snap_b[1:], snap_p[1:])) and \
        snap_b[0].as_tuple() == snap_p[0].as_tuple()
    for sim in sims:
        sim.close()
    vb, vp = n_particles * steps / wall_b, n_particles * plain_steps / wall_p
    vh = n_particles * plain_steps / wall_head          # batched, over the one-by-one form's steps
    if not emit:
        return {"batched": round(vb, 1), "one_by_one": round(vp, 1), "ratio": round(vh / vp, 2), "steps": steps,
                "one_by_one_steps": plain_steps,
                "instances": len(sizes), "particles": n_particles,
                "smallest_instance_same_alone_and_batched": bool(same)}
    print(json.dumps({
        "metric": "particle-steps/s", "workload": "c4", "value": round(vb, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic", "batched": round(vb, 1), "one_by_one": round(vp, 1), "ratio": round(vh / vp, 2),
        "batched_over_one_by_one_steps": round(vh, 1), "one_by_one_steps": plain_steps,
        "config": {"workload": "256 cells, C1 type mix at 5e-4 per nm^3, sizes log-uniform on [4096, 524288] "
                               "rescaled to 2^24 particles (tests/golden/c4_sizes.json), periodic, "
                               "C/R/D = 16/32/8", "instances": len(sizes), "particles": n_particles},
        "smallest_instance_same_alone_and_batched": bool(same)}))


def serving(args):
    """Requests of one C1-mix cell each arrive as a Poisson process (client
    threads submit them at their arrival times); the serving loop batches the
    resident ones.  Latency = completion time - arrival time."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(42)
    n = args.requests
    arrive = np.cumsum(rng.exponential(1.0 / args.rate, n))
    lo, hi = 8192, args.max_size
    s_lo, s_hi = (int(x) for x in args.request_steps.split(","))
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
    steps = rng.integers(s_lo, s_hi + 1, n)
    docs = [scenes.c1_mix_scene(int(sizes[i]), seed=1000 + i) for i in range(n)]
    tmpl = scenes.c1_mix_scene(8192, seed=0)
    budget = 200e9
    server_doc = {"hbm_budget_bytes": budget, "max_instances": 128, "max_cells": 1 << 27}
    srv = sc.Server(tmpl[0], tmpl[1], server_doc)
    t0 = time.perf_counter()
    ids = [None] * n

    def submit(i):
        wait = arrive[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids[i] = srv.submit(docs[i][0], docs[i][1], int(steps[i]))

    with ThreadPoolExecutor(8) as ex:
        list(ex.map(submit, range(n)))
    print(f"[c5] {n} requests submitted by {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    infos = [None] * n
    small = int(np.argmin(sizes * steps))
    pending = set(range(n))
    reported = 0
    while pending:
        for i in sorted(pending):
            info = srv.poll(ids[i])
            if info.state in (sc.REQ_DONE, sc.REQ_FAILED):
                if info.state == sc.REQ_FAILED:
                    raise RuntimeError(f"request {i} failed")
                infos[i] = info
                pending.discard(i)
                if i != small:
                    srv.release(ids[i])              # final states are not needed here
        done = n - len(pending)
        if done >= reported + 25 or (done == n and reported < n):
            reported = done
            print(f"[c5] {done} done at {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        if pending:
            time.sleep(0.05)
    got = srv.result(ids[small])
    sstats = srv.stats()
    srv.stop()
    sim = sc.Simulation(*docs[small])
    sim.step(int(steps[small]))
    want = sim.snapshot()
    sim.close()
    same = got[0].as_tuple() == want[0].as_tuple() and got[1].tobytes() == want[1].tobytes() and \
        got[2].tobytes() == want[2].tobytes()
    lat = np.array([infos[i].t_done - arrive[i] for i in range(n)])
    queue = np.array([infos[i].t_admit - arrive[i] for i in range(n)])
    work = float(np.sum(sizes * steps))
    span = max(x.t_done for x in infos) - float(arrive[0])
    offered = work / float(arrive[-1] - arrive[0]) if n > 1 else 0.0
    print(json.dumps({
        "metric": "request latency", "workload": "c5", "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "p50_s": round(float(np.percentile(lat, 50)), 3), "p90_s": round(float(np.percentile(lat, 90)), 3),
        "p99_s": round(float(np.percentile(lat, 99)), 3), "mean_queue_s": round(float(queue.mean()), 3),
        "throughput": round(work / span, 1), "offered": round(offered, 1), "unit_throughput": "particle-steps/s",
        "requests": n, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Poisson {args.rate}/s, {n} requests, C1-mix cells log-uniform on [{lo}, {hi}] "
                               f"particles, steps uniform on [{s_lo}, {s_hi}], C/R/D = 16/32/8, "
                               f"{budget / 1e9:.0f} GB budget", "max_instances": 128},
        "scheduler": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in sstats.items()},
        "smallest_request_same_as_alone": bool(same)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--box", type=float, default=1000.0)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="leave out the 256-cell batch comparison")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default scheduling options")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="library option for the simulation (sc_set_option), e.g. relax_split=1; "
                         "given options are applied after the autotuner")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5"], default="c3",
                    help="c3: 2^23 particles with a lipid shell and a dense protein ball (the headline line); "
                         "c2: one 2^20-particle uniform cell; c1: 32k-particle cell; c4: 256 cells batched vs "
                         "one by one; c5: serving loop under Poisson arrivals")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="c4: steps timed for the one-cell-at-a-time form (default: --steps)")
    ap.add_argument("--requests", type=int, default=300, help="c5: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="c5: arrivals per second")
    ap.add_argument("--max-size", type=int, default=524288, help="c5: largest request (particles; sizes log-uniform from 8192)")
    ap.add_argument("--request-steps", default="50,500", metavar="LO,HI", help="c5: steps per request, uniform on [LO, HI]")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with a plain copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload in ("c4", "c5"):
        (batched_vs_plain if args.workload == "c4" else serving)(args)
        return

    if args.workload == "c1":
        cfg, setup = scenes.c1_scene()
        workload = "C1: periodic 400 nm box, 8192 lipids, 8192 proteins, 2048 dummies, 14336 metabolite carriers"
    elif args.workload == "c3":
        cfg, setup = scenes.c3_scene()
        workload = ("C3: periodic 2000 nm box, lipid shell r 900 +- 5 (2097152), protein ball r 150 (1258291), "
                    "uniform 629146 dummies and 4404019 metabolite carriers")
    else:
        cfg, setup = scenes.uniform_scene(args.particles, args.box, boundary="periodic", init=True)
        workload = None
    t_init = time.perf_counter()
    sim = sc.Simulation(cfg, setup)
    sim.synchronize()
    t_init = time.perf_counter() - t_init
    P = sim.params
    for _ in range(args.warmup):                   # one at a time: list pools settle before the timing
        sim.step(1)
        sim.synchronize()
    print(f"[bench] init {t_init:.2f}s, warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    tune = None
    if not args.no_autotune:
        t = time.perf_counter()
        tune = sim.autotune(2)
        print(f"[bench] autotune {time.perf_counter() - t:.1f}s: {tune['chosen']}", file=sys.stderr, flush=True)
    for opt in args.option:
        name, value = opt.split("=")
        sim.set_option(name, int(value))
    sim.step(1)                                    # the chosen options' step graph, captured before the timing
    sim.synchronize()

    c_start = sim.counts()
    series = None
    t0 = time.perf_counter()
    if args.workload == "c3":                      # the dense ball relaxes: keep the per-step times
        series = [round(sim.time_steps(1), 3) for _ in range(args.steps)]
        ms_event = sum(series) / len(series)
    else:
        ms_event = sim.time_steps(args.steps)      # HIP events on the library's stream
    sim.synchronize()
    wall = time.perf_counter() - t0
    c_end = sim.counts()
    n_particles = c_end.n_active + c_end.n_active_metabolic
    assert c_end.step - c_start.step == args.steps
    value = n_particles * args.steps / wall
    print(f"[bench] {args.steps} steps in {wall:.3f}s ({ms_event:.3f} ms/step by events)", file=sys.stderr, flush=True)

    # dominant kernel pair: the relax pass (R per step; with relax_split one
    # launch of k_relax_pass and one of k_relax_coll), its start and end
    # recorded in HIP events of its own on the library's stream: within 3
    # more steps right after the timed ones (the pass among the step's other
    # kernels, as in the timed steps), and launches of one later pass back to
    # back (for comparison)
    loop_launch_ms = sim.time_kernel(sc.KERNEL_RELAX_PASS, 200 if args.workload != "c3" else 10)
    relax_launch_ms = sim.time_step_kernel(sc.KERNEL_RELAX_PASS, 3)

    # the last step's pair lists: entries, longest list, particles that walk the stencil instead
    lists = {}
    for q, names in ((0, ("relax", "relax_partners")), (1, ("relax_metabolic", "diffuse"))):
        cnt = sim.export(sc.EXPORT_LIST_COUNTS, q).reshape(-1, 2)
        for col, name in enumerate(names):
            walk = cnt[:, col] == sc.LIST_WALKS
            live = cnt[~walk, col]
            lists[name] = {"entries": int(live.sum()), "max": int(live.max(initial=0)), "walkers": int(walk.sum())}
    fp, fm = sim.fused_counts()       # relaxed in the fused kernel in the last step
    lists["fused"] = {"plain": fp, "metabolic": fm}

    # per-stage timing (eager launches bracketed by HIP events)
    prof = sim.profile(args.profile_steps)
    n_p, n_m = c_end.n_active, c_end.n_active_metabolic
    pops = setup["populations"]
    n_l = sum(p["count"] for p in pops if p["array"] == "plain" and p["type"] == 0)   # lipids (none are removed)
    passes = radix_passes(P)
    bytes_ = stage_bytes(n_p, n_l, n_m, passes, P)
    stages = {}
    for k, b in bytes_.items():
        ms = prof.get(k, 0.0)
        stages[k] = {"ms": round(ms, 4), "gbs": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    b_step = sum(bytes_.values())
    relax_bytes = 108 * n_p                        # 80 B read + 28 B written per plain particle
    achieved = relax_bytes / (relax_launch_ms * 1e-3) / 1e9
    # HBM bytes per relax pass from the counters of a profiled run of this
    # workload (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size)
    traffic = None
    # and its VALU instructions per launch (SQ_INSTS_VALU, wave-level) from
    # the same counter runs
    valu_insts = None
    tfile = os.path.join(ROOT, "profiles", f"relax_pass_traffic_{args.workload}.json")
    if (workload is not None or args.workload == "c2") and os.path.exists(tfile):
        with open(tfile) as f:
            tjson = json.load(f)
        traffic = tjson.get("hbm_bytes_per_launch")
        # the counted kernels are the ones this run launches only with the
        # same relax kernel options; the count scales with the list entries
        chosen = dict(tune["chosen"]) if tune else {"relax_split": 0, "fuse_bonds": 0}
        for opt in args.option:
            k, v = opt.split("=")
            chosen[k] = int(v)
        counted = tjson.get("options", {})
        same = all(chosen.get(k) == counted.get(k) for k in ("relax_split", "fuse_bonds"))
        if same and tjson.get("valu_insts_per_dispatch") and tjson.get("relax_entries"):
            valu_insts = tjson["valu_insts_per_dispatch"] * lists["relax"]["entries"] / tjson["relax_entries"]
    split = bool(tune["chosen"]["relax_split"]) if tune else False
    fused = bool(tune["chosen"].get("fuse_bonds", 0)) if tune else False   # the pass kernel with its bond terms
    for opt in args.option:
        if opt.startswith("relax_split="):
            split = opt.endswith("=1")
        if opt.startswith("fuse_bonds="):
            fused = opt.endswith("=1")

    out = {
        "metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (workload or f"single cell, {n_particles} particles, periodic {P.sim_size:.0f} nm "
                                           f"box, cutoff = cell = {P.grid_cell_size:.0f} nm, 20/20/55/5 mix") +
                               f", C/R/D = {P.noise_coordination_steps}/{P.relaxation_steps}/"
                               f"{P.metabolite_diffusion_steps}",
                   "particles": n_particles, "plain": n_p, "lipids": n_l, "metabolic": n_m,
                   "box_nm": P.sim_size, "graph": True},
        "hbm_roofline_step": {"bytes_per_step": b_step,
                              "frac": round(b_step * value / n_particles / (HBM_PEAK_GBS * 1e9), 4)},
        "roofline": {"bound": "hbm", "kernel": ("k_relax_pass_bt" if fused else "k_relax_pass") +
                                               (" + k_relax_coll" if split else ""),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launch_ms": round(relax_launch_ms, 5), "bytes_per_launch": relax_bytes,
                     "launch_ms_back_to_back": round(loop_launch_ms, 5)},
        "stages": stages,
        "lists": lists,
    }
    if valu_insts:
        # the pass at C3's density is bound by its arithmetic: VALU lane
        # instructions issued per second against the VALU's issue rate
        # (256 CUs x 4 SIMDs x 16 lanes per clock at 2.4 GHz; 157.3 TFLOP/s
        # fp32 is this x 2 for FMA x 2 for packed pairs)
        lane_rate = valu_insts * 64 / (relax_launch_ms * 1e-3) / 1e12
        out["valu_roofline"] = {"bound": "valu", "kernel": out["roofline"]["kernel"],
                                "achieved": round(lane_rate, 2), "peak": VALU_PEAK_TLANES,
                                "unit": "T VALU lane-instructions/s", "frac": round(lane_rate / VALU_PEAK_TLANES, 4),
                                "valu_insts_per_launch": int(valu_insts),
                                "counted": "SQ_INSTS_VALU of the counter run in profiles/, scaled by this run's "
                                           "relax list entries"}
    if tune is not None:
        out["autotune"] = tune
    if series is not None:
        out["ms_per_step_series"] = series
    sim.close()
    if not args.no_cpu:
        libs = native_oracles()
        wl = args.workload if workload is not None else "c2"
        out["cpu_baseline"] = cpu_baseline(wl, libs[0])
        out["cpu_baseline_threaded"] = cpu_baseline(wl, libs[1], threaded=True)
        print("[bench] cpu baselines done", file=sys.stderr, flush=True)
    if args.workload == "c3" and not args.no_c4:
        out["c4"] = batched_vs_plain(args, steps=5, warmup=2, emit=False)
        print("[bench] c4 done", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
