"""The physics extensions on the GPU against the oracle, bit for bit:
Stokes-Einstein Brownian noise (config noiseModel, an extension of
spatial_cell_sim/src/dynamics.cuh:156-215) and site-contact bonds between
two-site proteins (rows with bondProbability, an extension of
interactions.cuh:216-346) and metabolite reactions (setup key "reactions";
the reference only diffuses, metabolites.cuh:352-431), through the init
sequence and whole steps, staged, concurrent and batched."""
import numpy as np
import pytest

import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu


def _same(label, got, want):
    scenes.assert_same_state(label, *got, *want)


def _bond_pairs(c, p):
    """(id, partner) of every bond, and whether each is mirrored."""
    n = c.n_active
    bonds = {}
    for i in range(n):
        bonds[int(p["id"][i])] = {(int(b[0]), int(b[2])) for b in p["bonds"][i][:p["n_bonds"][i]]}
    mirrored = all((row, pid) in bonds.get(partner, set()) for pid, bs in bonds.items() for row, partner in bs
                   if row in (2001, 2002))
    return sum(len(bs) for bs in bonds.values()), mirrored


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_stokes_einstein_noise_matches_oracle(boundary):
    cfg, setup = scenes.mixed_scene(seed=31, boundary=boundary, init=True)
    cfg.update(scenes.STOKES_EINSTEIN)
    print(f"[stokes-einstein {boundary}]", flush=True)
    sim = sc.Simulation(cfg, setup)
    sim.step(3)
    got = sim.snapshot()
    sim.close()
    _same(f"stokes-einstein {boundary}", got, scenes.oracle_run(cfg, setup, 3))


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_site_bonds_match_oracle(boundary):
    cfg, setup = scenes.site_scene(boundary=boundary)
    print(f"[site bonds {boundary}] {cfg['numParticles']} plain", flush=True)
    sim = sc.Simulation(cfg, setup)
    init = sim.snapshot()
    _same("site bonds after init", init, scenes.oracle_run(cfg, setup, 0))
    sim.step(4)
    got = sim.snapshot()
    sim.close()
    _same(f"site bonds {boundary}, 4 steps", got, scenes.oracle_run(cfg, setup, 4))
    n0, _ = _bond_pairs(init[0], init[1])
    n, mirrored = _bond_pairs(got[0], got[1])
    assert n > n0 and n >= 20 and mirrored, (n0, n)


@pytest.mark.parametrize("boundary,n,L,contact", [
    ("periodic", 2400, 40.0, 3.0),
    ("clamp", 2400, 40.0, 3.0),
    ("periodic", 1200, 30.0, None),  # 3 cells per axis
])
def test_site_bonds_crowded_match_oracle(boundary, n, L, contact):
    """Crowded binding sites: with a 3 nm contact distance and every contact
    bonding (probability 1) a protein meets many candidates, so the site
    walk's per-lane queues fill and drain several times within one stencil
    walk; a 30 nm periodic box has 3 cells per axis."""
    cfg, setup = scenes.site_scene(n=n, L=L, boundary=boundary, prob=1.0, contact=contact, seed=17)
    print(f"[site crowded {boundary} n={n} L={L}]", flush=True)
    sim = sc.Simulation(cfg, setup)
    init = sim.snapshot()
    _same("crowded sites after init", init, scenes.oracle_run(cfg, setup, 0))
    sim.step(3)
    got = sim.snapshot()
    sim.close()
    _same(f"crowded sites {boundary}, 3 steps", got, scenes.oracle_run(cfg, setup, 3))
    n0, _ = _bond_pairs(init[0], init[1])
    nb, mirrored = _bond_pairs(got[0], got[1])
    assert nb > n0 and mirrored, (n0, nb)


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_reactions_match_oracle(boundary):
    cfg, setup = scenes.reaction_scene(boundary=boundary)
    print(f"[reactions {boundary}] stage by stage", flush=True)
    want = scenes.oracle_run(cfg, setup, 3)
    sim = sc.Simulation(cfg, setup)
    sim.step(3)
    got = sim.snapshot()
    sim.close()
    _same(f"reactions {boundary}", got, want)
    c, _, m = got
    met = m["metabolites"][:c.n_active_metabolic]
    assert (met[:, 1] > 0).sum() > 50 and (met[:, 2] > 0).sum() > 20, "reactions did not fire"
    print(f"[reactions {boundary}] concurrent", flush=True)
    sim = sc.Simulation(cfg, setup)
    sim.set_option("concurrent", 1)
    sim.step(3)
    got = sim.snapshot()
    sim.close()
    _same(f"reactions {boundary}, concurrent", got, want)


def test_reactions_batched_match_oracle():
    docs = [scenes.reaction_scene(seed=13 + k, n=600 + 300 * k) for k in range(3)]
    print("[reactions batch] 3 instances", flush=True)
    b = sc.Batch([c for c, _ in docs], [s for _, s in docs])
    b.step(2)
    for k in range(3):
        _same(f"reactions batch instance {k}", b.instance_snapshot(k), scenes.oracle_run(*docs[k], 2))
    b.close()
