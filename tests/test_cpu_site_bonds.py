"""Site-contact bonds (interaction rows with bondProbability, an extension
of the threshold predicates of spatial_cell_sim/src/interactions.cuh:319-333)
pinned on hand-placed proteins by what the rule states: the probability 1
and -> 0 limits, the contact radius of 1 nm around the first partner's site,
the lowest-id choice among competing candidates with bonds only between
partners that chose each other, the periodic minimum image, a group already
taken, two sites of one protein, and over many contacts bonds that always
come in pairs and form at the row's probability.  Planted mistakes (no
mutual check, the highest instead of the lowest key, no minimum image, a
draw that differs between the two partners) each fail one of these."""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc
from test_cpu_oracle import make_particles

SITE = [0.0, 0.0, 8.0]


def rows(prob, two_sites=False, contact=1.0):
    r = [{"id": 2001, "group": 2001, "firstPartnerType": 200, "secondPartnerType": 200,
          "relativeOrientation": [0, 0, 0, 1], "relativePosition": SITE,
          "bondProbability": prob, "contactDistance": contact}]
    if two_sites:
        r.append(dict(r[0], id=2002, group=2002, relativePosition=[0.0, 0.0, -8.0]))
    return r


def run(plain_rows, interactions, L=60.0, boundary="clamp", plant=None, steps=0):
    for i, r in enumerate(plain_rows):
        r.setdefault("id", i)
        r.setdefault("type", 200)
        r.setdefault("radius", 4.0)
    cfg = scenes.config(len(plain_rows) + 4, 4, L, boundary=boundary, cutoff=10.0)
    s = sc.Setup(cfg, {"types": scenes.TYPES, "interactions": interactions})
    plain = make_particles(plain_rows, cfg["numParticles"])
    met = make_particles([], 4)
    o = orc.Oracle(s.params, s.tables, sc.Counts(len(plain_rows), 0, len(plain_rows), 0), plain, met, plant=plant)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_COMPLEXIFY, 0)
    c, p, _ = o.snapshot()
    o.close()
    return {int(p["id"][i]): [tuple(b) for b in p["bonds"][i][:p["n_bonds"][i]]] for i in range(c.n_active)}


def check_limits_and_radius(plant=None):
    A = {"pos": [20.0, 20.0, 20.0]}
    for off, want in ((0.9, True), (1.1, False)):
        got = run([A, {"pos": [20.0 + off, 20.0, 28.0]}], rows(1.0), plant=plant)
        assert (got[0] == [(2001, 2001, 1)] and got[1] == [(2001, 2001, 0)]) == want, (off, got)
    # the p -> 0 limit (a probability of 0 selects the reference's threshold
    # predicate instead); the smallest draw is 2^-24
    got = run([A, {"pos": [20.3, 20.0, 28.0]}], rows(1e-9), plant=plant)
    assert got == {0: [], 1: []}


def check_conflict(plant=None):
    A = {"pos": [20.0, 20.0, 20.0]}
    B = {"pos": [20.3, 20.0, 28.0]}
    C = {"pos": [19.7, 20.0, 28.0]}
    got = run([A, B, C], rows(1.0), plant=plant)
    assert got[0] == [(2001, 2001, 1)] and got[1] == [(2001, 2001, 0)] and got[2] == [], got


def check_periodic(plant=None):
    # A's site lies across the x = L face; B sits just inside x = 0
    q = [0.0, np.sin(np.pi / 4), 0.0, np.cos(np.pi / 4)]          # rot_y(90 deg): the site points along +x
    A = {"pos": [55.0, 30.0, 30.0], "rot": q}
    B = {"pos": [2.5, 30.0, 30.0]}
    got = run([A, B], rows(1.0), L=60.0, boundary="periodic", plant=plant)
    assert got[0] == [(2001, 2001, 1)] and got[1] == [(2001, 2001, 0)], got


def check_group_and_two_sites(plant=None):
    A = {"pos": [20.0, 20.0, 20.0], "bonds": [[2001, 2001, 7]]}
    B = {"pos": [20.3, 20.0, 28.0]}
    got = run([A, B], rows(1.0), plant=plant)
    assert got[0] == [(2001, 2001, 7)] and got[1] == [], got
    A = {"pos": [20.0, 20.0, 20.0]}
    B = {"pos": [20.3, 20.0, 28.0]}
    C = {"pos": [19.7, 20.0, 12.0]}
    got = run([A, B, C], rows(1.0, two_sites=True), plant=plant)
    assert sorted(got[0]) == [(2001, 2001, 1), (2002, 2002, 2)], got
    assert got[1] == [(2001, 2001, 0)] and got[2] == [(2002, 2002, 0)], got


def check_pairs_and_rate(plant=None, prob=0.5):
    """Many separated pairs in contact: every bond is mirrored, and the
    fraction bonded is the row's probability (binomial, 4 sigma)."""
    rng = np.random.default_rng(3)
    prs = []
    for k in range(300):
        x, y, z = 15.0 + 30.0 * (k % 10), 15.0 + 30.0 * ((k // 10) % 10), 15.0 + 30.0 * (k // 100)
        prs.append({"pos": [x, y, z], "id": 2 * k})
        prs.append({"pos": [x + rng.uniform(-0.5, 0.5), y + rng.uniform(-0.5, 0.5), z + 8.0], "id": 2 * k + 1})
    got = run(prs, rows(prob), L=300.0, plant=plant)
    bonded = 0
    for k in range(300):
        a, b = got[2 * k], got[2 * k + 1]
        assert (a == [(2001, 2001, 2 * k + 1)]) == (b == [(2001, 2001, 2 * k)]), (k, a, b)
        bonded += a != []
    sd = np.sqrt(300 * prob * (1 - prob))
    assert abs(bonded - 300 * prob) < 4 * sd, bonded


def test_site_bond_limits_and_contact_radius():
    check_limits_and_radius()


def test_site_bond_lowest_id_and_mutual_choice():
    check_conflict()


def test_site_bond_periodic_minimum_image():
    check_periodic()


def test_site_bond_group_taken_and_two_sites():
    check_group_and_two_sites()


def test_site_bonds_paired_at_their_probability():
    check_pairs_and_rate()


PLANTS = [(51, check_conflict), (52, check_conflict), (53, check_periodic), (55, check_pairs_and_rate)]


@pytest.mark.parametrize("plant,check", PLANTS, ids=[str(p) for p, _ in PLANTS])
def test_planted_site_mistakes_fail(plant, check):
    check(plant=0)
    with pytest.raises(AssertionError):
        check(plant=plant)
