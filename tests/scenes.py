"""Scene documents (config + setup) shared by the tests, the smoke check and
the benchmark, plus bit-exact state comparison helpers."""
import gzip
import json
import os

import numpy as np

TYPES = [
    {"type": 0, "radius": 2.5, "hydrophobic": 2},   # lipid (POLAR)
    {"type": 1, "radius": 3.0},                     # dummy molecule
    {"type": 100, "radius": 2.0},                   # A: binds B, polymerises M
    {"type": 101, "radius": 2.0},                   # B
    {"type": 102, "radius": 1.5},                   # M: monomer
    {"type": 200, "radius": 4.0},                   # protein
    {"type": 1000, "radius": 1.0},                  # metabolite carrier
]

# Interaction table in the reference's complexification.json format that walks
# through bond formation, deferred and immediate transitions, state changes,
# polymerisation (start, extension, end) and break-after-transition.
REACTIONS = [
    {"id": 1001, "group": 1001, "firstPartnerType": 100, "secondPartnerType": 101,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 3.0],
     "orientationAlignmentThreshold": -2.0, "positionAlignmentThreshold": -1.0, "transitionTo": 1002},
    {"id": 1002, "group": 1001, "firstPartnerType": 100, "firstPartnerState": 1, "secondPartnerType": 101,
     "onlyViaTransition": True, "waitForAlignment": False, "transitionTo": 1003, "polymerize": 1004,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 3.0]},
    {"id": 1003, "group": 1001, "firstPartnerType": 100, "secondPartnerType": 101, "onlyViaTransition": True,
     "breakAfterTransition": True, "waitForState": False, "waitForAlignment": False,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 3.0]},
    {"id": 1004, "group": 1004, "firstPartnerType": 100, "secondPartnerType": 102, "secondPartnerState": 2,
     "onlyViaTransition": True, "relativeOrientation": [0, 0.3826834, 0, 0.9238795],
     "relativePosition": [0, 2.5, 0]},
    {"id": 1005, "group": 1004, "firstPartnerType": 100, "secondPartnerType": -1, "onlyViaTransition": True,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 0]},
    {"id": 1011, "group": 1011, "firstPartnerType": 100, "firstPartnerState": 1, "secondPartnerType": 101,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, -3.0],
     "orientationAlignmentThreshold": -2.0, "positionAlignmentThreshold": -1.0, "transitionTo": 1012},
    {"id": 1012, "group": 1011, "firstPartnerType": 100, "firstPartnerState": 2, "secondPartnerType": 101,
     "onlyViaTransition": True, "waitForAlignment": False, "polymerize": 1005,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, -3.0]},
    {"id": 1020, "group": 1020, "firstPartnerType": 200, "secondPartnerType": 200,
     "relativeOrientation": [0.7071068, 0, 0, 0.7071068], "relativePosition": [0, 0, 8.0],
     "orientationAlignmentThreshold": 0.0, "positionAlignmentThreshold": 0.3},
]


# The benchmark's protein (type 200): two binding sites, +-z at 8 nm (two
# touching r = 4 nm proteins), each bonding a partner whose centre comes
# within 1 nm of it with probability 0.1 per step (site-contact rows, the
# bondProbability extension of the interaction table).
SITE_ROWS = [
    {"id": 2001, "group": 2001, "firstPartnerType": 200, "secondPartnerType": 200,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0.0, 0.0, 8.0],
     "bondProbability": 0.1, "contactDistance": 1.0},
    {"id": 2002, "group": 2002, "firstPartnerType": 200, "secondPartnerType": 200,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0.0, 0.0, -8.0],
     "bondProbability": 0.1, "contactDistance": 1.0},
]

# The benchmark's dynamics: Stokes-Einstein diffusion at 310 K in water
# (0.69 mPa s) with a 0.1 ns timestep.
STOKES_EINSTEIN = {"noiseModel": "stokes_einstein", "temperature": 310.0, "viscosity": 6.9e-4, "timestep": 0.1}

# Metabolite reactions: one uncatalysed conversion, one catalysed by proteins
# (type 200) within 6 nm and one by lipids (type 0) within the 10 nm cell.
METABOLITE_REACTIONS = [
    {"from": 0, "to": 1, "rate": 0.05},
    {"from": 1, "to": 2, "rate": 0.02, "catalystType": 200, "catalystDistance": 6.0},
    {"from": 2, "to": 0, "rate": 0.01, "catalystType": 0, "catalystDistance": 10.0},
]
# metabolite 0 at 100 in the carriers of the box's lower fifth along x
SEED = [{"index": 0, "value": 100.0, "x_below_fraction": 0.2}]


def config(n_plain, n_met, L, boundary="clamp", seed=42, cell=10.0, cutoff=8.0, diff=10.0, C=16, R=32, D=8):
    return {
        "numParticles": n_plain, "numMetabolicParticles": n_met, "simSize": L,
        "gridCellSize": cell, "maxInteractionDistance": cutoff, "maxDiffusionDistance": diff,
        "movementNoiseScale": 0.7, "rotationNoiseScale": 0.05, "metaboliteMovementNoiseScale": 1.5,
        "velocityDecay": 0.7, "angularVelocityDecay": 0.7,
        "noiseCoordinationSteps": C, "relaxationSteps": R, "metaboliteDiffusionSteps": D,
        "persistEveryNthFrame": 1, "reportEveryNthFrame": 10, "steps": 10,
        "boundary": boundary, "seed": seed,
    }


def mixed_scene(scale=1.0, L=60.0, boundary="clamp", seed=42, init=False, **kw):
    """Every code path: lipids (hydrophobic alignment), type-1 and type-0
    chains, A/B/M reaction partners, self-binding proteins, metabolites."""
    def n(x):
        return max(1, int(x * scale))
    pops = [
        {"array": "plain", "type": 0, "count": n(300)},
        {"array": "plain", "type": 200, "count": n(60)},
        {"array": "plain", "type": 1, "count": n(10) * 8, "shape": "chains", "chain_length": 8, "spacing": 5.0},
        {"array": "plain", "type": 1, "count": n(4) * 4, "shape": "chains", "chain_length": 4, "spacing": 5.5,
         "bond_type": 0},
        {"array": "plain", "type": 100, "count": n(120)},
        {"array": "plain", "type": 101, "count": n(160)},
        {"array": "metabolic", "type": 1000, "count": n(500)},
    ]
    n_plain = sum(p["count"] for p in pops if p["array"] == "plain")
    n_met = sum(p["count"] for p in pops if p["array"] == "metabolic")
    cfg = config(n_plain + n(400), n_met, L, boundary=boundary, seed=seed, **kw)
    setup = {"types": TYPES, "interactions": REACTIONS, "populations": pops, "init": init,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.2},
                                 {"index": 3, "value": 7.5, "x_below_fraction": 0.5}]}
    return cfg, setup


def benchmark_physics(cfg, setup):
    """The dynamics of the benchmark configs: Stokes-Einstein noise at 310 K
    in water (timestep 0.1 ns), proteins (type 200) with two binding sites that
    bond with probability 0.1 per site contact within 1 nm, and the metabolite
    reactions of METABOLITE_REACTIONS."""
    cfg = dict(cfg, **STOKES_EINSTEIN)
    setup = dict(setup, interactions=list(setup.get("interactions", [])) + [dict(r) for r in SITE_ROWS],
                 reactions=METABOLITE_REACTIONS)
    return cfg, setup


def c1_scene(boundary="periodic", seed=42, init=True, physics=True):
    """Config 1 of the benchmark suite: 400 nm box, 10 nm cutoff = cell,
    8192 lipids (r 2.5), 8192 proteins (r 4), 2048 dummies (r 3), 14336
    metabolite carriers (r 1), uniform positions, orientations uniform on S3."""
    pops = [
        {"array": "plain", "type": 0, "count": 8192},
        {"array": "plain", "type": 200, "count": 8192},
        {"array": "plain", "type": 1, "count": 2048},
        {"array": "metabolic", "type": 1000, "count": 14336},
    ]
    cfg = config(18432, 14336, 400.0, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    setup = {"types": TYPES, "populations": pops, "init": init,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.2}]}
    return benchmark_physics(cfg, setup) if physics else (cfg, setup)


def site_scene(n=1200, L=60.0, boundary="periodic", seed=5, init=True, prob=0.1, noise=True, contact=None):
    """Proteins with two binding sites crowded enough (about one within a
    site's contact radius per few sites) that site bonds form every step,
    lipids and metabolite carriers around them (contact: another contact
    distance for both rows)."""
    rows = [dict(r, bondProbability=prob) for r in SITE_ROWS]
    if contact is not None:
        rows = [dict(r, contactDistance=contact) for r in rows]
    pops = [{"array": "plain", "type": 200, "count": n},
            {"array": "plain", "type": 0, "count": n // 4},
            {"array": "metabolic", "type": 1000, "count": n // 2}]
    cfg = config(n + n // 4, n // 2, L, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    if noise:
        cfg.update(STOKES_EINSTEIN)
    setup = {"types": TYPES, "interactions": rows, "populations": pops, "init": init}
    return cfg, setup


def reaction_scene(boundary="periodic", seed=13, n=1200, L=60.0):
    """site_scene with the metabolite carriers seeded and METABOLITE_REACTIONS
    on: every row fires (catalysts are within reach of most carriers)."""
    cfg, setup = site_scene(n=n, L=L, boundary=boundary, seed=seed)
    setup = dict(setup, reactions=METABOLITE_REACTIONS, metabolite_seed=[dict(SEED[0], x_below_fraction=0.6)])
    return cfg, setup


def dense_scene(boundary="clamp", seed=9, init=False):
    """Crowded: a protein ball, a lipid shell around it and a dense metabolite
    ball, so that many particles have more neighbour-list entries than a list
    holds in its slots (pool entries, stencil-walk fallback, pool growth)."""
    pops = [
        {"array": "plain", "type": 200, "count": 500, "shape": "ball", "ball_radius": 18.0},
        {"array": "plain", "type": 0, "count": 400, "shape": "shell", "shell_radius": 25.0, "thickness": 4.0},
        {"array": "metabolic", "type": 1000, "count": 800, "shape": "ball", "ball_radius": 22.0},
    ]
    cfg = config(1000, 900, 60.0, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    setup = {"types": TYPES, "populations": pops, "init": init}
    return cfg, setup


def c3_scene(boundary="periodic", seed=42, init=True, physics=True):
    """Config 3 of the benchmark suite: 2000 nm box, 2^23 particles: a lipid
    shell of radius 900 +- 5 nm about the centre (2,097,152), a protein ball of
    radius 150 nm (1,258,291), and uniform metabolite carriers : dummies = 7 : 1
    (4,404,019 : 629,146)."""
    pops = [
        {"array": "plain", "type": 0, "count": 2097152, "shape": "shell", "shell_radius": 900.0, "thickness": 10.0},
        {"array": "plain", "type": 200, "count": 1258291, "shape": "ball", "ball_radius": 150.0},
        {"array": "plain", "type": 1, "count": 629146},
        {"array": "metabolic", "type": 1000, "count": 4404019},
    ]
    n_plain = 2097152 + 1258291 + 629146
    cfg = config(n_plain, 4404019, 2000.0, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    setup = {"types": TYPES, "populations": pops, "init": init,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.2}]}
    return benchmark_physics(cfg, setup) if physics else (cfg, setup)


def c3_like_scene(f, boundary="periodic", seed=42, init=True, physics=True):
    """Config 3's structure with lengths scaled by f at the same densities:
    box 2000 f nm, lipid shell of radius 900 f and the same 10 nm thickness
    (2,097,152 f^2 lipids), protein ball of radius 150 f (1,258,291 f^3),
    uniform metabolite carriers : dummies = 7 : 1 (5,033,165 f^3 in all).
    f = 1 is c3_scene; f = 0.1 is a 200 nm cell of about 27k particles that
    the oracle steps in seconds."""
    n_shell = int(round(2097152 * f * f))
    n_ball = int(round(1258291 * f ** 3))
    n_uni = int(round(5033165 * f ** 3))
    n_dummy = n_uni // 8
    n_met = n_uni - n_dummy
    L = round(2000.0 * f / 10.0) * 10.0
    pops = [
        {"array": "plain", "type": 0, "count": n_shell, "shape": "shell", "shell_radius": 900.0 * f,
         "thickness": 10.0},
        {"array": "plain", "type": 200, "count": n_ball, "shape": "ball", "ball_radius": 150.0 * f},
        {"array": "plain", "type": 1, "count": n_dummy},
        {"array": "metabolic", "type": 1000, "count": n_met},
    ]
    n_plain = n_shell + n_ball + n_dummy
    cfg = config(n_plain, n_met, L, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    setup = {"types": TYPES, "populations": pops, "init": init,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.2}]}
    return benchmark_physics(cfg, setup) if physics else (cfg, setup)


def c4_sizes(n=256, total=1 << 24, seed=42):
    """Config 4 of the benchmark suite: n instance sizes drawn log-uniform on
    [4096, 524288] (numpy default_rng(seed)), rescaled so that they sum to
    exactly `total` (2^24) particles."""
    rng = np.random.default_rng(seed)
    raw = np.exp(rng.uniform(np.log(4096.0), np.log(524288.0), n))
    sizes = np.maximum(np.round(raw * (total / raw.sum())).astype(np.int64), 64)
    sizes[int(np.argmax(sizes))] += total - int(sizes.sum())
    return [int(x) for x in sizes]


def c1_mix_scene(n_total, seed, boundary="periodic", init=True, physics=True):
    """A cell with C1's type mix (8192 : 8192 : 2048 : 14336 lipids : proteins :
    dummies : metabolite carriers) and density 5e-4 per nm^3: box
    (n / 5e-4)^(1/3) nm rounded to whole 10 nm cells (at least 3)."""
    n_l = n_total * 8192 // 32768
    n_p = n_total * 8192 // 32768
    n_d = n_total * 2048 // 32768
    n_m = n_total - n_l - n_p - n_d
    L = max(30.0, 10.0 * round((n_total / 5e-4) ** (1.0 / 3.0) / 10.0))
    pops = [{"array": "plain", "type": 0, "count": n_l}, {"array": "plain", "type": 200, "count": n_p},
            {"array": "plain", "type": 1, "count": n_d}, {"array": "metabolic", "type": 1000, "count": n_m}]
    cfg = config(n_l + n_p + n_d, n_m, L, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    setup = {"types": TYPES, "populations": pops, "init": init,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.2}]}
    return benchmark_physics(cfg, setup) if physics else (cfg, setup)


def uniform_scene(n_total, L, boundary="periodic", seed=42, init=True, mix=(0.20, 0.20, 0.55, 0.05), physics=True):
    """Single cell with the 20/20/55/5 lipid/protein/metabolite/dummy mix."""
    n_l = int(n_total * mix[0])
    n_p = int(n_total * mix[1])
    n_d = int(n_total * mix[3])
    n_m = n_total - n_l - n_p - n_d
    pops = [
        {"array": "plain", "type": 0, "count": n_l},
        {"array": "plain", "type": 200, "count": n_p},
        {"array": "plain", "type": 1, "count": n_d},
        {"array": "metabolic", "type": 1000, "count": n_m},
    ]
    cfg = config(n_l + n_p + n_d, n_m, L, boundary=boundary, seed=seed, cutoff=10.0, diff=10.0)
    setup = {"types": TYPES, "populations": pops, "init": init,
             "metabolite_seed": [{"index": 0, "value": 100.0, "x_below_fraction": 0.2}]}
    return benchmark_physics(cfg, setup) if physics else (cfg, setup)


def oracle_run(cfg, setup, steps):
    """The oracle's state after creating the cell from its documents (with
    the init sequence of main.cu:283-338 unless the setup sets "init":
    false, as sc_create does) and `steps` steps: (counts, plain, metabolic)."""
    import oracle as orc
    import spatial_cell as sc
    s = sc.Setup(cfg, setup)
    o = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic)
    if setup.get("init", True):
        for stage, arg in ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1)):
            o.run_stage(stage, arg)
    o.step(steps)
    out = o.snapshot()
    o.close()
    return out


FLOAT_FIELDS = ["radius", "pos", "rot", "vel", "pos_noise", "ang_vel", "ang_noise", "debug", "metabolites"]
INT_FIELDS = ["id", "type", "flags", "state", "last_state_change_step", "hydrophobic", "n_bonds",
              "scheduled_polymerization", "bonds"]


def diff_states(a, b, n=None):
    """Bitwise comparison of two particle arrays (NaN equals NaN).  Returns a
    list of (field, first differing slot, count)."""
    if n is not None:
        a, b = a[:n], b[:n]
    out = []
    for f in INT_FIELDS:
        x, y = a[f], b[f]
        bad = np.nonzero((x != y).reshape(len(x), -1).any(axis=1))[0]
        if len(bad):
            out.append((f, int(bad[0]), len(bad)))
    for f in FLOAT_FIELDS:
        x = np.ascontiguousarray(a[f]).view(np.uint32)
        y = np.ascontiguousarray(b[f]).view(np.uint32)
        xf, yf = a[f], b[f]
        ne = (x != y) & ~(np.isnan(xf) & np.isnan(yf))
        bad = np.nonzero(ne.reshape(len(x), -1).any(axis=1))[0]
        if len(bad):
            out.append((f, int(bad[0]), len(bad)))
    return out


def assert_same_state(label, ca, pa, ma, cb, pb, mb):
    assert ca.as_tuple() == cb.as_tuple(), f"{label}: counts {ca.as_tuple()} != {cb.as_tuple()}"
    dp = diff_states(pa, pb)
    dm = diff_states(ma, mb)
    assert not dp and not dm, f"{label}: plain {dp} metabolic {dm}"


_HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE_SETUP = os.path.join(_HERE, "..", "spatial-cell_amd", "scenes", "reference_setup.json")


def reference_tables():
    """The reference's generated type, interaction and complex tables
    (fixture written by golden/make_reference_tables.py)."""
    with gzip.open(os.path.join(_HERE, "golden", "reference_tables.json.gz")) as f:
        return json.load(f)


def reference_scene(n_plain=65536, n_met=65536, L=400.0, seed=42, init=True):
    """The reference's hard-coded scene (setup/setup.cuh:18-253) as the
    package's setup document, with the tables inline instead of the
    ../universe-config/gen paths, at the reference config.json's sizes."""
    with open(REFERENCE_SETUP) as f:
        doc = json.load(f)
    for k in ("types_file", "interactions_file", "complexes_file"):
        doc.pop(k)
    doc.update(reference_tables())
    doc["init"] = init
    cfg = config(n_plain, n_met, L, boundary="clamp", seed=seed, cutoff=8.0, diff=10.0)
    return cfg, doc
