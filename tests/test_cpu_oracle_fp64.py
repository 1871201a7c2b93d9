"""The float oracle (which the HIP path equals bit for bit) against the same
steps in double precision (liboracle_f64.so: every real quantity of the step a
double, the C library's scalar functions; random draws and constants are the
same float values), over 100 steps.

Interacting particles make the step chaotic: a pair at a collision or
alignment threshold, or two nearly antiparallel polar axes, turn a rounding
difference into a different trajectory within a few steps, so bounds on
every particle hold only where particles do not interact.  Bounds stated
here:
- particles that never come within the interaction or diffusion distance
  (a 100 nm lattice): every position within 2e-3 nm (2e-6 of the box),
  every orientation within 1e-3 degrees, metabolites identical;
- a cell with the C1 type mix, density and benchmark physics (4,096
  particles), per particle: after 1 step 99 % of positions within 1e-4 nm,
  all within 0.05 nm, 99 % of orientations within 1e-2 degrees, at most
  0.5 % of them past 0.1 degree, and those only polar lipids with a polar
  neighbour inside the alignment distance 1.4 (r_p + r_t) (the rotation
  that takes -up to the hydrophobic direction, dynamics.cuh:805-815, is
  ill-conditioned where the particle's up axis points nearly along that
  direction); after 10 steps 90 % of positions within
  1e-4 nm, 99 % within 5e-3 nm, at most 1 % past 0.1 nm, 90 % of
  orientations within 1e-3 degrees and 99 % within 0.1 degree; metabolic
  particles within 2e-3 nm throughout;
- an interacting periodic scene (lipids, chains, reaction partners,
  metabolites): ensemble statistics -- mean squared displacement within
  20 %, bond count within 15 %, metabolite totals within 1e-3 relative,
  active particle count within 2 %.
Measured: lattice 6.8e-4 nm and 1.2e-4 degrees; C1 cell after 1 step 4.6e-5 /
2.1e-2 nm, 8.0e-4 degrees, 0.26 % past 0.1 degree, after 10 steps 6.0e-5 /
7.8e-4 nm, 0.13 % past 0.1 nm, 5.6e-5 / 3.2e-2 degrees, metabolic 7.2e-4 nm;
scene 7.6 % / 6 % / 1.5e-4 / equal."""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc
from test_cpu_oracle import make_particles


def _lattice(n_side, a, off, met, rng):
    rows = []
    for ix in range(n_side):
        for iy in range(n_side):
            for iz in range(n_side):
                q = rng.normal(size=4)
                q /= np.linalg.norm(q)
                r = {"id": len(rows) + (100000 if met else 0), "type": 1000 if met else [0, 1, 200][len(rows) % 3],
                     "radius": 1.0 if met else 3.0, "pos": [off + a * ix, off + a * iy, off + a * iz],
                     "rot": q.tolist()}
                if met:
                    r["metabolites"] = (rng.random(10) * 50).tolist()
                rows.append(r)
    return rows


def _by_id(a, n):
    a = a[:n]
    return a[np.argsort(a["id"])]


def _min_image(d, L):
    return d - L * np.round(d / L)


def test_fp32_and_fp64_agree_where_particles_do_not_interact():
    rng = np.random.default_rng(7)
    L, steps = 1000.0, 100
    # 20 nm of travel in 100 steps at most: sites 100 nm apart (plain) and
    # 86.6 nm from the metabolic ones never come within 10 nm of each other
    P = _lattice(10, 100.0, 5.0, False, rng)
    M = _lattice(10, 100.0, 55.0, True, rng)
    cfg = scenes.config(len(P) + 8, len(M) + 8, L, boundary="periodic")
    s = sc.Setup(cfg, {"types": scenes.TYPES, "interactions": []})
    out = []
    for fp64 in (False, True):
        o = orc.Oracle(s.params, s.tables, sc.Counts(len(P), len(M), 200000, 0),
                       make_particles(P, cfg["numParticles"]), make_particles(M, cfg["numMetabolicParticles"]),
                       fp64=fp64)
        o.step(steps)
        out.append(o.snapshot())
    (ca, pa, ma), (cb, pb, mb) = out
    assert ca.n_active == cb.n_active == len(P) and ca.n_active_metabolic == cb.n_active_metabolic == len(M)
    A, B = _by_id(pa, len(P)), _by_id(pb, len(P))
    assert np.array_equal(A["id"], B["id"])
    start = np.array([r["pos"] for r in P])
    travel = np.linalg.norm(_min_image(A["pos"] - start, L), axis=1)
    assert travel.max() < 30.0, "the lattice must stay non-interacting"
    d = np.linalg.norm(_min_image(A["pos"].astype(np.float64) - B["pos"], L), axis=1)
    assert d.max() <= 2e-3
    assert d.max() > 0.0, "the two builds must differ in their rounding"
    ra, rb = A["rot"].astype(np.float64), B["rot"].astype(np.float64)
    q = np.abs(np.sum(ra * rb, axis=1)) / (np.linalg.norm(ra, axis=1) * np.linalg.norm(rb, axis=1))
    ang = 2.0 * np.degrees(np.arccos(np.clip(q, 0.0, 1.0)))
    print("lattice: position", d.max(), "nm, orientation", ang.max(), "degrees")
    assert ang.max() <= 1e-3
    MA, MB = _by_id(ma, len(M)), _by_id(mb, len(M))
    dm = np.linalg.norm(_min_image(MA["pos"].astype(np.float64) - MB["pos"], L), axis=1)
    assert dm.max() <= 2e-3
    assert np.array_equal(MA["metabolites"], MB["metabolites"])


def _orientation_deg(ra, rb):
    ra, rb = ra.astype(np.float64), rb.astype(np.float64)
    q = np.abs(np.sum(ra * rb, axis=1)) / (np.linalg.norm(ra, axis=1) * np.linalg.norm(rb, axis=1))
    return 2.0 * np.degrees(np.arccos(np.clip(q, 0.0, 1.0)))


def test_fp32_and_fp64_per_particle_bounds_on_a_c1_cell():
    cfg, setup = scenes.c1_mix_scene(4096, 7, init=False)
    s = sc.Setup(cfg, setup)
    L = cfg["simSize"]
    pair = [orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic, fp64=f) for f in (False, True)]
    for o in pair:
        for st, a in ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1)):
            o.run_stage(st, a)
    c0, p0, _ = pair[0].snapshot()
    P0 = _by_id(p0, c0.n_active)
    done = 0
    for k in (1, 10):
        for o in pair:
            o.step(k - done)
        done = k
        (ca, pa, ma), (cb, pb, mb) = [o.snapshot() for o in pair]
        A, B = _by_id(pa, ca.n_active), _by_id(pb, cb.n_active)
        assert np.array_equal(A["id"], B["id"]) and np.array_equal(A["id"], P0["id"])
        d = np.linalg.norm(_min_image(A["pos"].astype(np.float64) - B["pos"], L), axis=1)
        ang = _orientation_deg(A["rot"], B["rot"])
        MA, MB = _by_id(ma, ca.n_active_metabolic), _by_id(mb, cb.n_active_metabolic)
        assert np.array_equal(MA["id"], MB["id"])
        dm = np.linalg.norm(_min_image(MA["pos"].astype(np.float64) - MB["pos"], L), axis=1)
        print(f"step {k}: position q90/q99/max {np.quantile(d, 0.9):.2e} {np.quantile(d, 0.99):.2e} {d.max():.2e} nm, "
              f"orientation q90/q99 {np.quantile(ang, 0.9):.2e} {np.quantile(ang, 0.99):.2e} deg, "
              f"past 0.1 deg {np.mean(ang > 0.1):.4f}, metabolic max {dm.max():.2e} nm")
        assert d.max() > 0.0, "the two builds must differ in their rounding"
        assert dm.max() <= 2e-3
        if k == 1:
            assert np.quantile(d, 0.99) <= 1e-4 and d.max() <= 0.05
            assert np.quantile(ang, 0.99) <= 1e-2 and np.mean(ang > 0.1) <= 0.005
            pos0 = P0["pos"].astype(np.float64)
            polar = P0["hydrophobic"] == sc.POLAR
            for i in np.where(ang > 0.1)[0]:
                assert polar[i], "an orientation outlier that is not polar"
                dist = np.linalg.norm(_min_image(pos0 - pos0[i], L), axis=1)
                reach = 1.4 * (P0["radius"].astype(np.float64) + float(P0["radius"][i]))
                near = polar & (dist <= reach)
                near[i] = False
                assert near.any(), "an orientation outlier without an alignment partner"
        else:
            assert np.quantile(d, 0.9) <= 1e-4 and np.quantile(d, 0.99) <= 5e-3 and np.mean(d > 0.1) <= 0.01
            assert np.quantile(ang, 0.9) <= 1e-3 and np.quantile(ang, 0.99) <= 0.1


def _ensemble(fp64, cfg, setup):
    s = sc.Setup(cfg, setup)
    o = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic, fp64=fp64)
    for st, a in ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1)):
        o.run_stage(st, a)
    c0, p0, _ = o.snapshot()
    o.step(100)
    c, p, m = o.snapshot()
    L = cfg["simSize"]
    P0 = _by_id(p0, c0.n_active)
    P = p[:c.n_active]
    P = P[np.isin(P["id"], P0["id"])]
    P = P[np.argsort(P["id"])]
    d = _min_image(P["pos"].astype(np.float64) - P0["pos"], L)
    act = (p[:c.n_active]["flags"] & sc.FLAG_ACTIVE) != 0
    M = m[:c.n_active_metabolic]
    mact = (M["flags"] & sc.FLAG_ACTIVE) != 0
    return {"msd": float(np.mean(np.sum(d * d, axis=1))),
            "bonds": int(p[:c.n_active]["n_bonds"][act].sum()),
            "metabolites": M["metabolites"][mact].astype(np.float64).sum(axis=0),
            "active": int(c.n_active)}


def test_fp32_and_fp64_ensembles_agree_after_100_steps():
    cfg, setup = scenes.mixed_scene(scale=0.5, L=50.0, boundary="periodic", seed=22)
    a = _ensemble(False, cfg, setup)
    b = _ensemble(True, cfg, setup)
    print("fp32", a, "\nfp64", b)
    assert abs(a["msd"] / b["msd"] - 1.0) <= 0.20
    assert abs(a["bonds"] - b["bonds"]) <= 0.15 * b["bonds"] and b["bonds"] > 0
    rel = np.abs(a["metabolites"] - b["metabolites"]) / np.maximum(np.abs(b["metabolites"]), 1.0)
    assert rel.max() <= 1e-3
    assert abs(a["active"] - b["active"]) <= 0.02 * b["active"]


@pytest.mark.parametrize("fp64", [False, True])
def test_fp64_build_is_loaded_on_request(fp64):
    assert orc.load(fp64=fp64)._name.endswith("liboracle_f64.so" if fp64 else "liboracle.so")
