"""Metabolite reactions (setup key "reactions"; no reference counterpart: the
reference only diffuses metabolites, spatial_cell_sim/src/metabolites.cuh:352-431)
pinned by what the rule fixes:

* mass balance: a row moves metabolite from `from` to `to`, the sum of the
  ten concentrations holds to rounding;
* exponential decay: T steps of an uncatalysed row leave m0 * (1 - rate)^T;
* saturation: min(rate * n, 1) empties `from` in one step once rate * n >= 1;
* table order: each row acts on what the previous one left;
* the catalyst count: brute force over every plain particle (minimum image
  in a periodic box) of active catalysts of the row's type within the row's
  distance -- a catalyst inside the bounding cube but outside the sphere,
  another type, an inactive one, one across the periodic face;
* the parser's limits.

Planted mistakes in the oracle -- the cube test alone (61), no saturation
(62), metabolite lost instead of moved (63) -- each fail one of these."""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc
from test_cpu_oracle import make_particles

CATALYST = 200


def run(met_rows, plain_rows, reactions, L=60.0, boundary="clamp", plant=None, times=1, cell=10.0):
    cfg = scenes.config(len(plain_rows) + 4, len(met_rows) + 4, L, boundary=boundary, cell=cell, cutoff=cell)
    s = sc.Setup(cfg, {"types": scenes.TYPES, "interactions": [], "reactions": reactions})
    for i, r in enumerate(plain_rows):
        r.setdefault("id", i)
        r.setdefault("type", CATALYST)
        r.setdefault("radius", 4.0)
    for i, r in enumerate(met_rows):
        r.setdefault("id", 100000 + i)
        r.setdefault("type", 1000)
        r.setdefault("radius", 1.0)
    plain = make_particles(plain_rows, cfg["numParticles"])
    met = make_particles(met_rows, cfg["numMetabolicParticles"])
    o = orc.Oracle(s.params, s.tables, sc.Counts(len(plain_rows), len(met_rows), 200000, 0), plain, met, plant=plant)
    o.run_stage(sc.STAGE_SORT)
    for _ in range(times):
        o.run_stage(sc.STAGE_REACT)
    c, _, m = o.snapshot()
    o.close()
    return {int(m["id"][i]): m["metabolites"][i].astype(np.float64) for i in range(c.n_active_metabolic)}


def row(frm, to, rate, cat=-1, dist=None):
    r = {"from": frm, "to": to, "rate": rate, "catalystType": cat}
    if dist is not None:
        r["catalystDistance"] = dist
    return r


M0 = [1.0, 0.25, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, 0.125, 0.0]


def check_balance_and_decay(plant=None):
    got = run([{"pos": [30.0, 30.0, 30.0], "metabolites": M0}], [], [row(0, 4, 0.1)], plant=plant, times=40)
    m = got[100000]
    assert abs(m.sum() - sum(M0)) < 1e-5, m
    assert abs(m[0] - (0.9 ** 40)) < 1e-5 * 40, m[0]
    assert abs(m[4] - (1.0 - 0.9 ** 40)) < 1e-5 * 40, m[4]
    assert m[1] == 0.25 and m[3] == 0.5 and m[8] == 0.125


def check_saturation(plant=None):
    cats = [{"pos": [30.0 + dx, 30.0, 30.0]} for dx in (-3.0, 2.0, 4.0)]
    got = run([{"pos": [30.0, 30.0, 30.0], "metabolites": M0}], cats, [row(3, 5, 0.4, CATALYST, 5.0)], plant=plant)
    m = got[100000]
    assert m[3] == 0.0 and m[5] == 0.5, m            # rate * n = 1.2 -> the whole of m[3]


def check_order():
    got = run([{"pos": [30.0, 30.0, 30.0], "metabolites": M0}], [], [row(0, 6, 0.5), row(6, 7, 0.5)])
    m = got[100000]
    assert (m[0], m[6], m[7]) == (0.5, 0.25, 0.25), m


def brute_count(mpos, cats, L, dist, periodic):
    n = 0
    for c in cats:
        if c.get("flags", 1) == 0 or c.get("type", CATALYST) != CATALYST:
            continue
        d = np.asarray(c["pos"], np.float64) - np.asarray(mpos, np.float64)
        if periodic:
            d -= L * np.round(d / L)
        n += np.linalg.norm(d) <= dist
    return n


def check_catalyst_count(plant=None, boundary="periodic"):
    """Hand-placed cases first, then a random scene against brute force."""
    L, dist = 60.0, 3.0
    rate = 1.0 / 64.0                          # 1 - n / 64 is exact for the counts here
    hand = [
        {"pos": [31.5, 31.5, 31.5]},           # |d| = 2.60: counts
        {"pos": [28.0, 28.0, 28.0]},           # each |d_i| = 2 < 3 but |d| = 3.46: the cube is not the sphere
        {"pos": [30.0, 32.5, 30.0], "type": 1},            # another type
        {"pos": [30.0, 30.0, 27.5], "flags": 0},           # inactive
        {"pos": [32.9, 30.0, 30.0]},           # |d| = 2.9
    ]
    met = [{"pos": [30.0, 30.0, 30.0], "metabolites": [1.0] + [0.0] * 9}]
    if boundary == "periodic":
        hand.append({"pos": [58.6, 1.0, 30.0]})           # across the x and y faces from (0.5, 59.2, 30)
        met.append({"pos": [0.5, 59.2, 30.0], "metabolites": [1.0] + [0.0] * 9})
    got = run(met, hand, [row(0, 1, rate, CATALYST, dist)], L=L, boundary=boundary, plant=plant)
    want = [2, 1] if boundary == "periodic" else [2]
    for k, w in enumerate(want):
        assert (1.0 - got[100000 + k][0]) * 64.0 == w, (k, got[100000 + k])

    rng = np.random.default_rng(11)
    cats, mets = [], []
    periodic = boundary == "periodic"
    for _ in range(900):
        cats.append({"pos": list(rng.uniform(0.0, L, 3)), "type": CATALYST if rng.random() < 0.8 else 1,
                     "flags": 1 if rng.random() < 0.9 else 0})
    for _ in range(300):
        mets.append({"pos": list(rng.uniform(0.0, L, 3)), "metabolites": [1.0] + [0.0] * 9})
    # drop what lies within 1e-3 of the distance, where float and double rounding may disagree
    def near(mp):
        for c in cats:
            d = np.asarray(c["pos"]) - np.asarray(mp)
            if periodic:
                d -= L * np.round(d / L)
            if abs(np.linalg.norm(d) - dist) < 1e-3:
                return True
        return False
    mets = [m for m in mets if not near(m["pos"])]
    got = run(mets, [dict(c) for c in cats], [row(0, 1, rate, CATALYST, dist)], L=L, boundary=boundary, plant=plant)
    hits = 0
    for k, m in enumerate(mets):
        n = brute_count(m["pos"], cats, L, dist, periodic)
        hits += n > 0
        assert (1.0 - got[100000 + k][0]) * 64.0 == n, (k, n, got[100000 + k][0])
    assert hits > 50


def test_reaction_balance_and_exponential_decay():
    check_balance_and_decay()


def test_reaction_saturates_at_the_whole_pool():
    check_saturation()


def test_reactions_apply_in_table_order():
    check_order()


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_catalyst_count_brute_force(boundary):
    check_catalyst_count(boundary=boundary)


def test_inactive_metabolic_particles_do_not_react():
    got = run([{"pos": [30.0, 30.0, 30.0], "metabolites": M0, "flags": 0},
               {"pos": [40.0, 30.0, 30.0], "metabolites": M0}], [], [row(0, 1, 0.5)])
    assert list(got[100000]) == M0
    assert got[100001][0] == 0.5 and got[100001][1] == 0.75


@pytest.mark.parametrize("bad,msg", [
    ([row(0, 0, 0.1)], "distinct"),
    ([row(0, 10, 0.1)], "distinct"),
    ([row(0, 1, -0.1)], "rate"),
    ([row(0, 1, 0.1, CATALYST, 12.0)], "catalystDistance"),
    ([row(0, 1, 0.1)] * (sc.MAX_REACTIONS + 1), "at most"),
])
def test_reaction_rows_are_validated(bad, msg):
    cfg = scenes.config(8, 8, 60.0)
    with pytest.raises(sc.SpatialCellError, match=msg):
        sc.Setup(cfg, {"types": scenes.TYPES, "interactions": [], "reactions": bad})


PLANTS = [(61, check_catalyst_count), (62, check_saturation), (63, check_balance_and_decay)]


@pytest.mark.parametrize("plant,check", PLANTS, ids=[str(p) for p, _ in PLANTS])
def test_planted_reaction_mistakes_fail(plant, check):
    check(plant=0)
    with pytest.raises(AssertionError):
        check(plant=plant)


def test_reaction_table_round_trip():
    rx = [row(0, 1, 0.05), row(1, 2, 0.02, CATALYST, 6.0)]
    s = sc.Setup(scenes.config(8, 8, 60.0), {"types": scenes.TYPES, "interactions": [], "reactions": rx})
    t = s.tables["reactions"]
    assert list(t["from"]) == [0, 1] and list(t["to"]) == [1, 2] and list(t["catalyst_type"]) == [-1, CATALYST]
    assert np.allclose(t["rate"], [0.05, 0.02]) and t["catalyst_distance"][1] == 6.0
    assert t["catalyst_distance"][0] == 8.0           # default: maxInteractionDistance
