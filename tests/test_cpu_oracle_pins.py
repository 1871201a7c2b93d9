"""Pins three oracle pieces that whole-step parity alone cannot check, each
against an independent answer, and shows that each check fails on a
mistake planted in the oracle (the ORACLE_PLANTS build, orc_plant ids of
oracle/oracle.cpp):

* removeInterferingMetabolicParticles (spatial_cell_sim/src/metabolites.cuh:18-88):
  an O(N^2) brute force of "any active lipid within r_m + r_l + 0.001"
  (the sum rounded to float, as the reference's float local does), pairs
  exactly at and one ulp past the threshold, non-lipid and inactive
  neighbours that must be ignored, and the count of active metabolic
  particles that the reference does not decrement (main.cu:301-312).
* move (dynamics.cuh:19-154): debugVector.xyz set to zero on active
  particles, .w kept, inactive slots copied unchanged.
* the periodic boundary of applyVelocities (the extension of
  dynamics.cuh:264-270 to a periodic box): a particle leaving through each
  face re-enters through the opposite one, a position landing exactly on L
  maps to 0, a tiny negative position whose sum with L rounds to L maps to 0.
"""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc
from test_cpu_oracle import make_particles

F32 = np.float32


def _oracle(plain_rows, met_rows, L, boundary, plant, cap_extra=4, debug=None, **cfgkw):
    cfg = scenes.config(len(plain_rows) + cap_extra, len(met_rows) + cap_extra, L, boundary=boundary, **cfgkw)
    s = sc.Setup(cfg, {"types": scenes.TYPES, "interactions": []})
    plain = make_particles(plain_rows, cfg["numParticles"])
    met = make_particles(met_rows, cfg["numMetabolicParticles"])
    if debug is not None:
        for arr, rows in ((plain, plain_rows), (met, met_rows)):
            arr["debug"][:len(rows)] = debug
    counts = sc.Counts(len(plain_rows), len(met_rows), len(plain_rows) + len(met_rows), 0)
    return orc.Oracle(s.params, s.tables, counts, plain, met, plant=plant), s


# ---- removeInterferingMetabolicParticles ---------------------------------------

def _interference_scene(seed=3):
    """Random metabolic particles among active lipids, inactive lipids and
    other plain particles in x in [20, 60]; hand-placed pairs at x in [0, 5]
    with exactly representable separations along x (the metabolic particle at
    x = 0, so the float difference is the lipid's coordinate itself)."""
    rng = np.random.default_rng(seed)
    plain, met = [], []
    box = lambda n: rng.uniform([20, 2, 2], [60, 58, 58], size=(n, 3)).astype(F32)  # noqa: E731
    for p in box(150):
        plain.append({"pos": p, "type": 0, "radius": 2.5, "hydrophobic": 2})
    for p in box(40):
        plain.append({"pos": p, "type": 0, "radius": 2.5, "hydrophobic": 2, "flags": 0})
    for p in box(100):
        plain.append({"pos": p, "type": 1, "radius": 3.0})
    for p in box(300):
        met.append({"pos": p, "type": 1000, "radius": 1.0})
    for p in box(10):
        met.append({"pos": p, "type": 1000, "radius": 1.0, "flags": 0})
    thr = F32(np.float64(F32(1.0) + F32(2.5)) + 0.001)      # the reference's float interferenceDist
    cases = [
        ("lipid exactly at r_m + r_l + 0.001", 0, thr, True),
        ("lipid one ulp past the threshold", 0, np.nextafter(thr, F32(10)), False),
        ("lipid between r_m + r_l and the threshold", 0, F32(3.5005), True),
        ("lipid at r_m + r_l", 0, F32(3.5), True),
        ("inactive lipid well inside", 0, F32(2.0), False),
        ("dummy (type 1) well inside", 1, F32(2.0), False),
    ]
    expect = {}
    for k, (label, kind, dx, removed) in enumerate(cases):
        y = 5.0 + 9.0 * (k % 6)
        z = 5.0 + 24.0 * (k // 6)
        mid = len(met)
        met.append({"pos": [0.0, y, z], "type": 1000, "radius": 1.0})
        if label.startswith("inactive"):
            plain.append({"pos": [dx, y, z], "type": 0, "radius": 2.5, "hydrophobic": 2, "flags": 0})
        elif kind == 1:
            plain.append({"pos": [dx, y, z], "type": 1, "radius": 3.0})
        else:
            plain.append({"pos": [dx, y, z], "type": 0, "radius": 2.5, "hydrophobic": 2})
        expect[mid] = (label, removed)
    for i, r in enumerate(plain):
        r["id"] = i
    for i, r in enumerate(met):
        r["id"] = i
    return plain, met, expect


def _brute_force_removed(plain, met):
    """Metabolic rows that an active lipid lies within the float
    interference distance of (all pairs, float32 arithmetic of the
    difference, its squared length summed x, y, z and the square root)."""
    lp = np.array([r["pos"] for r in plain if r["type"] == 0 and r.get("flags", 1) & 1], dtype=F32)
    lr = np.array([r["radius"] for r in plain if r["type"] == 0 and r.get("flags", 1) & 1], dtype=F32)
    out = set()
    for i, r in enumerate(met):
        if not (r.get("flags", 1) & 1):
            continue
        p = np.array(r["pos"], dtype=F32)
        d = lp - p
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        thr = (np.float64(F32(r["radius"]) + lr) + 0.001).astype(F32)
        if np.any(dist <= thr):
            out.add(i)
    return out


def check_remove_interfering(plant=None):
    plain, met, expect = _interference_scene()
    o, _ = _oracle(plain, met, 60.0, "clamp", plant, cutoff=8.0)
    c0, p0, m0 = o.snapshot()
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_REMOVE_INTERFERING)
    c1, p1, m1 = o.snapshot()
    # active count not decremented (main.cu:301-312 only clears flags)
    assert c1.as_tuple() == c0.as_tuple()
    n = c1.n_active_metabolic
    active = {int(m1["id"][i]): bool(m1["flags"][i] & 1) for i in range(n)}
    assert len(active) == len(met)
    for mid, (label, removed) in expect.items():
        assert active[mid] == (not removed), label
    want = _brute_force_removed(plain, met)
    for i, r in enumerate(met):
        was = bool(r.get("flags", 1) & 1)
        assert active[i] == (was and i not in want), f"metabolic {i}"
    assert len(want) > 20                       # the random part removes a fair share
    # nothing else changes: the plain array is only sorted, metabolic
    # records keep every other field
    o2 = {int(p1["id"][i]): i for i in range(c1.n_active)}
    for i in range(c0.n_active):
        j = o2[int(p0["id"][i])]
        assert p0[i].tobytes() == p1[j].tobytes()
    pos1 = {int(m1["id"][i]): m1["pos"][i] for i in range(n)}
    for i in range(c0.n_active_metabolic):
        assert np.array_equal(pos1[int(m0["id"][i])], m0["pos"][i])


def test_remove_interfering_brute_force_and_threshold():
    check_remove_interfering()


# ---- move: debugVector reset ---------------------------------------------------

def check_move_debug_reset(plant=None):
    rng = np.random.default_rng(5)
    rows = []
    for k in range(24):
        rows.append({"pos": rng.uniform(5, 55, 3).astype(F32), "radius": 2.0, "type": 1,
                     "flags": 0 if k % 5 == 3 else 1, "vel": rng.normal(size=3).astype(F32)})
    mrows = [dict(r, type=1000, radius=1.0) for r in rows[:12]]
    dbg = np.array([1.5, -2.25, 3.0, 4.75], F32)
    o, _ = _oracle(rows, mrows, 60.0, "clamp", plant, debug=dbg)
    _, p0, m0 = o.snapshot()
    o.run_stage(sc.STAGE_MOVE)
    _, p1, m1 = o.snapshot()
    for before, after, src in ((p0, p1, rows), (m0, m1, mrows)):
        for i, r in enumerate(src):
            if r["flags"] & 1:
                assert np.array_equal(after["debug"][i], [0, 0, 0, dbg[3]]), i
            else:
                assert np.array_equal(after["debug"][i], dbg), i
                assert before[i].tobytes() == after[i].tobytes()
            for f in ("pos", "vel", "rot", "ang_vel", "id", "flags"):
                assert np.array_equal(before[f][i], after[f][i]), (f, i)


def test_move_resets_debug_xyz_keeps_w():
    check_move_debug_reset()


# ---- applyVelocities in a periodic box -------------------------------------------

def _landing(target, start, L):
    """A velocity v with float32(start + float32(v * 0.45)) == target exactly."""
    v = F32((target - start) / 0.45)
    for _ in range(64):
        s = F32(F32(start) + F32(v * F32(0.45)))
        if s == F32(target):
            return v
        v = np.nextafter(v, F32(np.inf) if s < target else F32(-np.inf))
    raise AssertionError("no velocity lands on the target")


def _wrap_expected(x, L):
    """Map a float32 coordinate into [0, L): one period added or removed in
    float32; a sum that rounds up to L is L itself, i.e. 0."""
    L = F32(L)
    if x >= L:
        x = F32(x - L)
    elif x < 0:
        x = F32(x + L)
    return F32(0.0) if x >= L else x


def check_periodic_wrap(plant=None):
    L = 60.0
    rows = []
    # each face crossed in each direction, an inside move, exact landings
    for axis in range(3):
        for start, v in ((58.5, 7.0), (1.5, -6.0), (30.0, 3.0)):
            p = [30.0, 30.0, 30.0]
            vel = [0.0, 0.0, 0.0]
            p[axis], vel[axis] = start, v
            rows.append({"pos": p, "vel": vel})
        p = [30.0, 30.0, 30.0]
        vel = [0.0, 0.0, 0.0]
        p[axis], vel[axis] = 59.0, _landing(60.0, 59.0, L)          # lands exactly on L
        rows.append({"pos": p, "vel": vel})
        p = [30.0, 30.0, 30.0]
        vel = [0.0, 0.0, 0.0]
        p[axis] = 1.0
        vel[axis] = np.nextafter(_landing(0.0, 1.0, L), F32(-np.inf))   # a hair below 0
        rows.append({"pos": p, "vel": vel})
    for i, r in enumerate(rows):
        r.update(id=i, radius=2.0, type=1)
    o, s = _oracle(rows, [], L, "periodic", plant)
    _, p0, _ = o.snapshot()
    o.run_stage(sc.STAGE_APPLY_VELOCITIES)
    _, p1, _ = o.snapshot()
    step = F32(0.45)
    decay = F32(s.params.velocity_decay)
    for i, r in enumerate(rows):
        x = (np.asarray(r["pos"], F32) + (np.asarray(r["vel"], F32) * step).astype(F32)).astype(F32)
        want = np.array([_wrap_expected(F32(c), L) for c in x], F32)
        assert np.array_equal(p1["pos"][i], want), (i, r["pos"], r["vel"], p1["pos"][i], want)
        assert np.all((p1["pos"][i] >= 0) & (p1["pos"][i] < L))
        assert np.array_equal(p1["vel"][i], (np.asarray(r["vel"], F32) * decay).astype(F32))
    # the landing cases really land on the faces before wrapping
    assert F32(F32(59.0) + F32(rows[3]["vel"][0] * step)) == F32(60.0)
    assert F32(F32(1.0) + F32(rows[4]["vel"][0] * step)) < 0


def test_periodic_wrap_every_face_and_exact_landings():
    check_periodic_wrap()


# ---- the checks fail on planted mistakes -----------------------------------------

PLANTS = [
    (11, check_remove_interfering, "the metabolic radius taken twice"),
    (12, check_remove_interfering, "any plain particle interferes, not only lipids"),
    (13, check_remove_interfering, "threshold without the 0.001"),
    (14, check_remove_interfering, "strict < at the threshold"),
    (21, check_periodic_wrap, "a position landing on L stays L"),
    (22, check_periodic_wrap, "z clamped instead of wrapped"),
    (31, check_move_debug_reset, "debugVector not reset"),
    (32, check_move_debug_reset, "debugVector.w reset too"),
    (33, check_move_debug_reset, "inactive slots reset too"),
]


@pytest.mark.parametrize("plant,check,what", PLANTS, ids=[str(p[0]) for p in PLANTS])
def test_planted_mistake_is_caught(plant, check, what):
    check(plant=0)                 # the planting build without a plant passes
    with pytest.raises(AssertionError):
        check(plant=plant)
