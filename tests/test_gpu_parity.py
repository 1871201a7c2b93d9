"""GPU-vs-oracle parity through the C ABI: every stage of a step, whole steps
replayed from the captured hipGraph, the binning exports, frames and
edge cases.  Integer, index and bond state must match bit for bit; floating
point state is compared bitwise too (the kernels keep the oracle's operation
order and rounding), with NaN equal to NaN."""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu


def pair(cfg, setup):
    """A GPU simulation (init skipped) and an oracle on the same raw state."""
    sim = sc.Simulation(cfg, setup)
    c, p, m = sim.snapshot()
    o = orc.Oracle(sim.params, sim.tables(), c, p, m)
    return sim, o


def check(label, sim, o):
    ca, pa, ma = sim.snapshot()
    cb, pb, mb = o.snapshot()
    scenes.assert_same_state(label, ca, pa, ma, cb, pb, mb)


def check_binning(label, sim, o):
    for pop in (0, 1):
        for what in (sc.EXPORT_KEYS, sc.EXPORT_PERMUTATION, sc.EXPORT_RANGES, sc.EXPORT_NEIGHBORS):
            a = sim.export(what, pop)
            b = o.export(what, pop)
            assert a.shape == b.shape and np.array_equal(a, b), f"{label}: export {what} pop {pop}"


def run_step_by_stage(sim, o, P, label, check_every=True):
    def both(stage, arg=0, tag=""):
        print(f"[{label}] stage {sc.STAGE_NAMES[stage]} {arg}", flush=True)
        sim.run_stage(stage, arg)
        o.run_stage(stage, arg)
        if check_every:
            check(f"{label}/{sc.STAGE_NAMES[stage]}{tag}", sim, o)
    step = sim.counts().step
    both(sc.STAGE_SORT)
    check_binning(f"{label}/sort", sim, o)
    both(sc.STAGE_MOVE)
    for j in range(P.noise_coordination_steps):
        both(sc.STAGE_COORDINATE, j, f"[{j}]")
    both(sc.STAGE_APPLY_NOISE)
    for j in range(P.relaxation_steps):
        both(sc.STAGE_RELAX, j, f"[{j}]")
        if j % 2 == 0:
            both(sc.STAGE_RELAX_METABOLIC, j, f"[{j}]")
            both(sc.STAGE_RELAX_PARTNERS, j, f"[{j}]")
    both(sc.STAGE_APPLY_VELOCITIES)
    both(sc.STAGE_COMPLEXIFY, step)
    both(sc.STAGE_TRANSITION, step)
    both(sc.STAGE_POLYMERIZE, step)
    for j in range(P.metabolite_diffusion_steps):
        both(sc.STAGE_DIFFUSE, j, f"[{j}]")
    both(sc.STAGE_END_STEP)


def init_both(sim, o):
    for stage, arg in ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1)):
        sim.run_stage(stage, arg)
        o.run_stage(stage, arg)
        check(f"init/{sc.STAGE_NAMES[stage]}", sim, o)


def test_primitives_device_equals_host():
    import ctypes
    lib = sc.load()
    rng = np.random.default_rng(7)
    xs = {
        0: np.concatenate([rng.uniform(-7, 7, 20000), np.linspace(-3.2, 3.2, 4001)]),
        1: np.concatenate([rng.uniform(-7, 7, 20000), np.linspace(-3.2, 3.2, 4001)]),
        2: np.concatenate([rng.uniform(-1, 1, 20000), [-1.0, 1.0, 0.5, -0.5, 0.0, 1.0000001, np.nan]]),
        3: rng.uniform(1e-7, 10, 20000), 4: rng.uniform(1e-7, 10, 20000), 5: rng.uniform(-30, 30, 20000),
        6: rng.uniform(1e-3, 1, 20000), 7: rng.uniform(0, 100, 20000), 8: rng.uniform(-70000, 70000, 20000),
        9: rng.uniform(-1e3, 1e3, 20000), 10: rng.uniform(-1e3, 1e3, 20000), 11: rng.uniform(-1e3, 1e3, 20000),
    }
    for fn, x in xs.items():
        x = np.ascontiguousarray(x, dtype=np.float32)
        yd = np.zeros_like(x)
        yh = np.zeros_like(x)
        assert lib.sc_eval_primitive(0, fn, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     yd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(x), 1) == 0
        assert lib.sc_eval_primitive(0, fn, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     yh.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(x), 0) == 0
        same = (yd.view(np.uint32) == yh.view(np.uint32)) | (np.isnan(yd) & np.isnan(yh))
        assert same.all(), f"primitive {fn}: {np.count_nonzero(~same)} differ, first x={x[~same][:3]}"


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_stage_parity_two_steps(boundary):
    cfg, setup = scenes.mixed_scene(boundary=boundary)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    for s in range(2):
        run_step_by_stage(sim, o, sim.params, f"{boundary}/step{s}")
    fa, fb = sim.emit_frame(), o.emit_frame()
    assert fa == fb


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_graph_steps_bit_exact(boundary):
    cfg, setup = scenes.mixed_scene(boundary=boundary, seed=3)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    for chunk in (1, 2, 5, 12):
        print(f"[graph {boundary}] {chunk} steps", flush=True)
        sim.step(chunk)
        o.step(chunk)
        check(f"{boundary}/graph+{chunk}", sim, o)
    check_binning(f"{boundary}/graph", sim, o)
    assert sim.emit_frame() == o.emit_frame()


@pytest.mark.parametrize("C,R,D,boundary", [(1, 1, 1, "clamp"), (3, 5, 3, "periodic"), (0, 2, 0, "periodic"),
                                            (2, 7, 2, "clamp")])
def test_pass_counts_other_than_16_32_8(C, R, D, boundary):
    """Other sub-iteration counts: an odd relax count leaves the velocity
    buffers at the other parity after every step (two step graphs take
    turns), counts below three relax or two metabolic passes never reach the
    shortened worklists, and zero coordination or diffusion passes leave those
    stages out of the step."""
    cfg, setup = scenes.mixed_scene(boundary=boundary, seed=5 + R, C=C, R=R, D=D)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    for chunk in (1, 2, 3):
        print(f"[passes {C}/{R}/{D} {boundary}] {chunk} steps", flush=True)
        sim.step(chunk)
        o.step(chunk)
        check(f"{C}/{R}/{D}/{boundary}+{chunk}", sim, o)
    assert sim.emit_frame() == o.emit_frame()


def test_c1_trajectory_100_steps():
    """The benchmark's 32k-particle cell (400 nm periodic box, cutoff = cell =
    10 nm, 8192 lipids, 8192 proteins, 2048 dummies, 14336 metabolite
    carriers, C/R/D = 16/32/8) from its init sequence for 100 steps: the HIP
    path equals the oracle bit for bit after steps 1, 10 and 100.  The oracle
    runs its OpenMP build, which equals the serial one slot for slot."""
    cfg, setup = scenes.c1_scene()
    sim = sc.Simulation(cfg, setup)
    c, p, m = sim.snapshot()
    o = orc.Oracle(sim.params, sim.tables(), c, p, m, threaded=True)
    done = 0
    for target in (1, 10, 100):
        print(f"[c1 trajectory] to step {target}", flush=True)
        sim.step(target - done)
        o.step(target - done)
        done = target
        check(f"c1+{target}", sim, o)
    assert sim.emit_frame() == o.emit_frame()
    sim.close()


def test_eager_equals_graph():
    cfg, setup = scenes.mixed_scene(seed=11)
    a = sc.Simulation(cfg, setup)
    b = sc.Simulation(cfg, setup)
    b.set_option("graph", 0)
    for _ in range(3):
        a.step(2)
        b.step(2)
        ca, pa, ma = a.snapshot()
        cb, pb, mb = b.snapshot()
        scenes.assert_same_state("eager-vs-graph", ca, pa, ma, cb, pb, mb)


def test_init_in_create_matches_oracle():
    cfg, setup = scenes.mixed_scene(seed=5)
    raw = sc.Simulation(cfg, setup)
    c, p, m = raw.snapshot()
    o = orc.Oracle(raw.params, raw.tables(), c, p, m)
    for stage, arg in ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1)):
        o.run_stage(stage, arg)
    setup = dict(setup, init=True)
    sim = sc.Simulation(cfg, setup)
    check("create+init", sim, o)
    sim.step(3)
    o.step(3)
    check("create+init+3", sim, o)


def test_reactions_happen_and_match():
    """The reaction table must actually exercise transitions, state changes and
    polymerisation within a few steps (otherwise parity proves little)."""
    cfg, setup = scenes.mixed_scene(seed=9, scale=1.5, L=55.0)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    c0 = sim.counts()
    sim.step(6)
    o.step(6)
    check("reactions+6", sim, o)
    c, p, m = sim.snapshot()
    n = c.n_active
    types = set()
    for k in range(8):
        types |= set(p["bonds"][:n, k, 0][p["n_bonds"][:n] > k].tolist())
    assert c.n_active > c0.n_active, "no monomer was polymerised"
    assert {1, 0}.issubset(types)
    assert any(t >= 1000 for t in types)
    assert (p["state"][:n] > 0).any()


def test_propagate_state_in_the_step():
    """With "propagateState" on, the step runs propagateState after
    polymerize (interactions.cuh:743-822); with the reaction table's bonds
    propagating state, states must spread along bonds and match the oracle
    bit for bit, stage by stage and over graph-replayed steps."""
    cfg, setup = scenes.mixed_scene(seed=9, scale=1.5, L=55.0)
    cfg = dict(cfg, propagateState=True)
    setup = dict(setup, interactions=[dict(r, propagateState=True) for r in setup["interactions"]])
    sim, o = pair(cfg, setup)
    assert sim.params.propagate_state == 1
    init_both(sim, o)
    sim.step(4)
    o.step(4)
    check("propagate+4", sim, o)
    for tag in ("a", "b"):
        sim.run_stage(sc.STAGE_PROPAGATE_STATE, 0)
        o.run_stage(sc.STAGE_PROPAGATE_STATE, 0)
        check(f"propagate stage {tag}", sim, o)
    sim.step(2)
    o.step(2)
    check("propagate+6", sim, o)
    c, p, m = sim.snapshot()
    assert (p["state"][:c.n_active] > 0).any()


def test_edge_cases_particle_on_box_face_and_one_cell():
    """Particles exactly at pos == L, all particles in one cell, empty cells,
    a count that is not a multiple of the block size."""
    cfg = scenes.config(517, 300, 40.0, cutoff=8.0)
    setup = {"types": scenes.TYPES, "init": False, "populations": [
        {"array": "plain", "type": 0, "count": 200, "shape": "ball", "ball_radius": 2.0, "center": [5, 5, 5]},
        {"array": "plain", "type": 1, "count": 117},
        {"array": "plain", "type": 200, "count": 200, "shape": "ball", "ball_radius": 30.0,
         "center": [40.0, 40.0, 40.0]},
        {"array": "metabolic", "type": 1000, "count": 300, "shape": "ball", "ball_radius": 3.0,
         "center": [40.0, 0.0, 40.0]}]}
    sim, o = pair(cfg, setup)
    c, p, m = sim.snapshot()
    assert (p["pos"][:517] == 40.0).any(), "no particle on the box face"
    init_both(sim, o)
    sim.step(2)
    o.step(2)
    check("edge+2", sim, o)
    check_binning("edge", sim, o)


@pytest.mark.parametrize("n_plain,n_met", [(300, 0), (0, 300)])
def test_empty_population(n_plain, n_met):
    """One of the two arrays holds no particle at all (capacity kept): every
    stage still matches the oracle, the empty array's included."""
    cfg = scenes.config(max(n_plain, 64), max(n_met, 64), 50.0, boundary="periodic", seed=31)
    pops = []
    if n_plain:
        pops += [{"array": "plain", "type": 0, "count": n_plain // 2},
                 {"array": "plain", "type": 200, "count": n_plain - n_plain // 2}]
    if n_met:
        pops += [{"array": "metabolic", "type": 1000, "count": n_met}]
    setup = {"types": scenes.TYPES, "interactions": scenes.REACTIONS, "init": False, "populations": pops,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.3}]}
    print(f"[empty population {n_plain} {n_met}]", flush=True)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    run_step_by_stage(sim, o, sim.params, f"empty{n_plain}/{n_met}", check_every=False)
    check(f"empty{n_plain}/{n_met}/stages", sim, o)
    sim.step(2)
    o.step(2)
    check(f"empty{n_plain}/{n_met}+2", sim, o)


def test_full_capacity_refuses_new_monomers_like_the_oracle():
    """Capacity equal to the particle count: polymerisation has no free slot,
    on the device exactly as in the oracle (interactions.cuh:558-740)."""
    rows = scenes.mixed_scene(scale=0.6, L=40.0, seed=33)
    cfg, setup = rows
    pops = setup["populations"]
    n_plain = sum(pp["count"] for pp in pops if pp["array"] == "plain")
    n_met = sum(pp["count"] for pp in pops if pp["array"] == "metabolic")
    cfg = dict(cfg, numParticles=n_plain, numMetabolicParticles=n_met)
    print(f"[full capacity {n_plain} {n_met}]", flush=True)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    for k in range(4):
        sim.step(1)
        o.step(1)
        check(f"full+{k + 1}", sim, o)
    assert sim.counts().n_active <= n_plain


def test_full_size_invariants():
    """The benchmark's 2^20-particle cell, where the oracle would take
    minutes: properties that do not depend on size.  Every particle id is
    still there exactly once, positions stay finite and inside the periodic
    box, and the run is deterministic: graph replay and kernel-by-kernel
    launches give the same state bit for bit."""
    cfg, setup = scenes.uniform_scene(1 << 20, 1000.0, boundary="periodic", init=True)
    print("[full size: graph]", flush=True)
    a = sc.Simulation(cfg, setup)
    c0, p0, m0 = a.snapshot()
    a.step(3)
    ca, pa, ma = a.snapshot()
    for before, after, n0, n1 in ((p0, pa, c0.n_active, ca.n_active), (m0, ma, c0.n_active_metabolic,
                                                                        ca.n_active_metabolic)):
        assert n1 == n0
        assert np.array_equal(np.sort(before["id"][:n0]), np.sort(after["id"][:n1]))
        pos = after["pos"][:n1]
        assert np.isfinite(pos).all() and (pos >= 0.0).all() and (pos < 1000.0).all()
        assert np.isfinite(after["vel"][:n1]).all()
    a.close()
    print("[full size: eager]", flush=True)
    b = sc.Simulation(cfg, setup)
    b.set_option("graph", 0)
    b.step(3)
    cb, pb, mb = b.snapshot()
    b.close()
    scenes.assert_same_state("full-size graph-vs-eager", ca, pa, ma, cb, pb, mb)


def test_restore_roundtrip_and_inactive_tail():
    cfg, setup = scenes.mixed_scene(seed=21)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    sim.step(1)
    o.step(1)
    c, p, m = sim.snapshot()
    # deactivate a few plain particles and restore on both engines
    p["flags"][[3, 50, 51]] = 0
    sim.restore(c, p, m)
    o2 = orc.Oracle(sim.params, sim.tables(), c, p, m)
    sim.step(2)
    o2.step(2)
    check("restore+2", sim, o2)


def test_time_kernel_leaves_state_unchanged():
    """sc_time_kernel runs later passes without swapping buffers: the
    current state (what a snapshot shows) must not change."""
    cfg, setup = scenes.mixed_scene(scale=0.5, L=40.0, boundary="periodic", seed=3, init=True)
    sim = sc.Simulation(cfg, setup)
    with pytest.raises(sc.SpatialCellError):
        sim.time_kernel(sc.KERNEL_RELAX_PASS, 2)          # no step yet: no neighbour lists
    sim.step(2)
    before = sim.snapshot()
    for k in (sc.KERNEL_RELAX_PASS, sc.KERNEL_METABOLIC_PASS, sc.KERNEL_DIFFUSE_PASS):
        assert sim.time_kernel(k, 3) > 0.0
    after = sim.snapshot()
    assert before[1].tobytes() == after[1].tobytes() and before[2].tobytes() == after[2].tobytes()
    sim.step(1)                                           # and the simulation carries on


def test_time_step_kernel_steps_like_step():
    """sc_time_step_kernel runs real steps (kernel by kernel, the relax
    passes with dispatch events): the state after them equals sc_step's,
    bit for bit, and the mean launch time is positive."""
    cfg, setup = scenes.mixed_scene(scale=0.5, L=40.0, boundary="periodic", seed=5, init=True)
    a = sc.Simulation(cfg, setup)
    b = sc.Simulation(cfg, setup)
    a.step(2)
    ms = a.time_step_kernel(sc.KERNEL_RELAX_PASS, 2)
    assert ms > 0.0
    b.step(4)
    ca, pa, ma = a.snapshot()
    cb, pb, mb = b.snapshot()
    assert ca.step == cb.step == 4
    scenes.assert_same_state("time_step_kernel-vs-step", ca, pa, ma, cb, pb, mb)
    with pytest.raises(sc.SpatialCellError):
        a.time_step_kernel(sc.KERNEL_DIFFUSE_PASS, 1)     # only the relax pass
    a.close()
    b.close()


def test_xcd_run_does_not_change_results():
    """The tile order of the neighbour kernels (option xcd_run) only moves
    work between XCDs: every value gives the same state bit for bit."""
    cfg, setup = scenes.c1_scene()
    states = []
    for run in (0, 1, 4, 32):
        print(f"[xcd_run {run}]", flush=True)
        sim = sc.Simulation(cfg, setup)
        sim.set_option("xcd_run", run)
        sim.step(3)
        c, p, m = sim.snapshot()
        states.append((run, p.tobytes(), m.tobytes()))
        sim.close()
    for run, p, m in states[1:]:
        assert p == states[0][1] and m == states[0][2], run


@pytest.mark.parametrize("scene", ["c1", "dense"])
def test_relax_split_and_gather_do_not_change_results(scene):
    """The relax pass as one kernel or as two (general, collision-only;
    option relax_split), the collision-only kernel's gather width (option
    coll_gather), the general pass kernel's register budget (option
    pass_waves), and the bond terms in a launch of their own or in the pass
    kernel (option fuse_bonds; the C1 cell's proteins bond) give the same
    state bit for bit, on a C1 cell and on a dense scene whose lists run into
    the pools (and equal the oracle there)."""
    cfg, setup = scenes.c1_scene() if scene == "c1" else scenes.dense_scene(init=True)
    states = []
    # (split, gather, fuse_bonds, pass_waves, graph): the split's two kernels
    # run as parallel graph branches, or on two streams when launched eagerly
    for split, g, fb, w, graph in ((0, 8, 0, 3, 1), (1, 8, 0, 3, 1), (1, 4, 0, 3, 1), (0, 8, 1, 3, 1),
                                   (1, 4, 1, 3, 1), (0, 8, 0, 4, 1), (1, 4, 0, 4, 1), (1, 8, 0, 3, 0)):
        print(f"[relax_split {split} coll_gather {g} fuse_bonds {fb} pass_waves {w} graph {graph} {scene}]",
              flush=True)
        sim = sc.Simulation(cfg, setup)
        sim.set_option("relax_split", split)
        sim.set_option("coll_gather", g)
        sim.set_option("fuse_bonds", fb)
        sim.set_option("pass_waves", w)
        sim.set_option("graph", graph)
        sim.step(3)
        states.append(sim.snapshot())
        sim.close()
    for st in states[1:]:
        scenes.assert_same_state(f"relax options {scene}", *states[0], *st)
    if scene == "dense":
        scenes.assert_same_state("relax options dense vs oracle", *states[2], *scenes.oracle_run(cfg, setup, 3))


def test_sort_tile_size_does_not_change_results():
    """Onesweep tiles of 1024, 2048 or 4096 keys (option sort_items 4, 8,
    16) give the same stable permutation, cell table and state: the 32k C1
    cell (8 to 32 tiles per pass) after 3 steps, and its binning exports
    against the oracle's after the first sort."""
    cfg, setup = scenes.c1_scene()
    states = []
    for items in (16, 8, 4):
        print(f"[sort_items {items}]", flush=True)
        if items == 4:
            b = sc.Simulation(cfg, setup)
            b.set_option("sort_items", items)
            c, p, m = b.snapshot()
            o = orc.Oracle(b.params, b.tables(), c, p, m)
            b.run_stage(sc.STAGE_SORT, 0)
            o.run_stage(sc.STAGE_SORT, 0)
            check_binning("sort_items 4", b, o)
            o.close()
            b.close()
        sim = sc.Simulation(cfg, setup)
        sim.set_option("sort_items", items)
        sim.step(3)
        states.append(sim.snapshot())
        sim.close()
    for st in states[1:]:
        scenes.assert_same_state("sort tile sizes", *states[0], *st)


def test_autotune_leaves_state_and_results_unchanged():
    """sc_autotune tries every candidate from a checkpoint and restores it:
    the state right after tuning is the state before, and the steps after
    tuning equal untuned steps bit for bit."""
    cfg, setup = scenes.mixed_scene(scale=2.0, L=90.0, seed=17, init=True)
    a = sc.Simulation(cfg, setup)
    a.step(2)
    before = a.snapshot()
    rep = a.autotune(2)
    print("[autotune]", rep, flush=True)
    assert len(rep["candidates"]) >= 4 and rep["chosen"]["relax_split"] in (0, 1)
    scenes.assert_same_state("autotune restores the state", *before, *a.snapshot())
    a.step(3)
    b = sc.Simulation(cfg, setup)
    b.step(5)
    scenes.assert_same_state("tuned vs untuned", *a.snapshot(), *b.snapshot())
    a.close()
    b.close()


def test_reference_scene_parity():
    """The reference's own scene and tables (setup/setup.cuh: the lipid
    half-layers, 1002 linked chains, two complexes of 89 particles, 648
    interaction rows): init and two steps, bit for bit."""
    cfg, setup = scenes.reference_scene(init=False)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    check("reference/init", sim, o)
    for s in range(2):
        print(f"[reference scene] step {s}", flush=True)
        sim.step(1)
        o.step(1)
        check(f"reference/step{s}", sim, o)
    c, p, m = sim.snapshot()
    assert (p["n_bonds"][:c.n_active] > 0).sum() > 8000          # the chains stay linked


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_fused_relaxation_matches_per_pass(boundary):
    """Whole steps with the fused relaxation of small components give the
    same state bit for bit as the per-pass kernels alone, and the fused
    kernel does take most particles of a C1 cell."""
    cfg, setup = scenes.c1_scene(boundary=boundary)
    states = []
    for fuse in (0, 1):
        print(f"[fuse {fuse} {boundary}]", flush=True)
        sim = sc.Simulation(cfg, setup)
        sim.set_option("fuse", fuse)
        sim.step(4)
        c, p, m = sim.snapshot()
        states.append((p.tobytes(), m.tobytes()))
        if fuse:
            fused = sim.fused_counts()
            assert fused[0] > 0.8 * c.n_active and fused[1] > 0.8 * c.n_active_metabolic, fused
        sim.close()
    assert states[0][0] == states[1][0] and states[0][1] == states[1][1]


def test_concurrent_schedule_matches_sequential():
    """Steps with the side stream (captured as a graph or launched eagerly)
    give the same state bit for bit as the one-stream schedule."""
    cfg, setup = scenes.c1_scene()
    states = []
    for concurrent, graph in ((0, 1), (1, 1), (2, 1), (2, 0)):
        print(f"[concurrent {concurrent} graph {graph}]", flush=True)
        sim = sc.Simulation(cfg, setup)
        sim.set_option("concurrent", concurrent)
        sim.set_option("graph", graph)
        sim.step(3)
        c, p, m = sim.snapshot()
        states.append(((concurrent, graph), p.tobytes(), m.tobytes()))
        sim.close()
    for key, p, m in states[1:]:
        assert p == states[0][1] and m == states[0][2], key


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_dense_scene_pool_and_walk_paths(boundary):
    """Particles with more list entries than slots: pool runs, the walk of
    the stencil when the pool is short, and the pool growing between steps
    must all give the oracle's results bit for bit."""
    cfg, setup = scenes.dense_scene(boundary=boundary)
    sim, o = pair(cfg, setup)
    init_both(sim, o)
    walkers = []
    for s in range(4):
        print(f"[dense {boundary}] step {s}", flush=True)
        sim.step(1)
        o.step(1)
        check(f"dense/{boundary}/step{s}", sim, o)
        counts = np.concatenate([sim.export(sc.EXPORT_LIST_COUNTS, q) for q in (0, 1)])
        assert counts.max(initial=0) > 16                 # some lists run into the pool
        walkers.append(int((counts == sc.LIST_WALKS).sum()))
    # the first step's pool is short; after it has grown, every list fits
    assert walkers[0] > 0 and walkers[-1] == 0, walkers
