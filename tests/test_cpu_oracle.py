"""Pins the oracle against what the mathematics fixes: closed forms of the
cell key and grid size (spatial_cell_sim/src/config.cuh:66-74, grid.cuh:13-25),
std::stable_sort-equivalent permutation (CUB SortPairs), brute-force cell
ranges and neighbour lists, the integrator of an isolated particle from the
Philox stream (dynamics.cuh:156-282), two-body collision and diffusion closed
forms (dynamics.cuh:775-786, metabolites.cuh:140-431), bond formation /
transition / polymerisation on hand-placed partners (interactions.cuh:216-740)
and the frames.dat byte layout (storage.cuh:22-39, types.cuh:238-294)."""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc

F32 = np.float32


def make_particles(rows, n_cap):
    """rows: dicts with id/type/pos/radius/... -> PARTICLE_DTYPE array of n_cap slots."""
    a = np.zeros(n_cap, dtype=sc.PARTICLE_DTYPE)
    for i, r in enumerate(rows):
        a[i]["id"] = r.get("id", i)
        a[i]["type"] = r.get("type", 1)
        a[i]["flags"] = r.get("flags", 1)
        a[i]["state"] = r.get("state", 0)
        a[i]["last_state_change_step"] = r.get("last_state_change_step", -1)
        a[i]["scheduled_polymerization"] = -1
        bonds = r.get("bonds", [])
        a[i]["n_bonds"] = len(bonds)
        for k, b in enumerate(bonds):
            a[i]["bonds"][k] = b
        a[i]["hydrophobic"] = r.get("hydrophobic", 0)
        a[i]["radius"] = r.get("radius", 3.0)
        a[i]["pos"] = r["pos"]
        a[i]["rot"] = r.get("rot", [0, 0, 0, 1])
        a[i]["vel"] = r.get("vel", [0, 0, 0])
        a[i]["ang_vel"] = r.get("ang_vel", [0, 0, 0, 1])
        a[i]["pos_noise"] = r.get("pos_noise", [0, 0, 0])
        a[i]["ang_noise"] = r.get("ang_noise", [0, 0, 0, 1])
        a[i]["metabolites"] = r.get("metabolites", [0] * 10)
    return a


def oracle_for(plain_rows, met_rows, L=60.0, boundary="clamp", interactions=None, cap_extra=8, **cfgkw):
    cfg = scenes.config(len(plain_rows) + cap_extra, len(met_rows) + cap_extra, L, boundary=boundary, **cfgkw)
    s = sc.Setup(cfg, {"types": scenes.TYPES, "interactions": interactions or []})
    plain = make_particles(plain_rows, cfg["numParticles"])
    met = make_particles(met_rows, cfg["numMetabolicParticles"])
    counts = sc.Counts(len(plain_rows), len(met_rows), len(plain_rows), 0)
    return orc.Oracle(s.params, s.tables, counts, plain, met), s


# ---- grid and keys -----------------------------------------------------------

@pytest.mark.parametrize("L,cell,n,bits", [(400.0, 10.0, 64, 6), (1000.0, 10.0, 128, 7), (2000.0, 10.0, 256, 8),
                                           (10.0, 10.0, 2, 1), (64.0, 1.0, 64, 6), (65.0, 1.0, 128, 7),
                                           (100.0, 10.0, 16, 4)])
def test_keys_closed_form(L, cell, n, bits):
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, L, (300, 3)).astype(F32)
    pos[0] = [L, L, L]
    pos[1] = [0, 0, 0]
    pos[2] = [cell, 2 * cell, L]
    rows = [{"pos": p} for p in pos]
    rows[5]["flags"] = 0
    o, _ = oracle_for(rows, [], L=L, cell=cell, cutoff=min(cell, 8.0), diff=min(cell, 10.0))
    o.run_stage(sc.STAGE_SORT)
    g = np.minimum(np.floor(pos / F32(cell)).astype(np.int64), n - 1)   # float32 divide, truncate, clamp
    key = (g[:, 0] << (2 * bits)) | (g[:, 1] << bits) | g[:, 2]
    key[5] = 0xFFFFFFFF
    perm = np.argsort(key, kind="stable")
    assert np.array_equal(o.export(sc.EXPORT_PERMUTATION, 0), perm.astype(np.int32))
    assert np.array_equal(o.export(sc.EXPORT_KEYS, 0), key[perm].astype(np.uint32))
    assert key[0] == ((min(int(L / cell), n - 1) << (2 * bits)) | (min(int(L / cell), n - 1) << bits)
                      | min(int(L / cell), n - 1))


def test_sort_is_stable_over_previous_order_and_ranges_brute_force():
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 30, (500, 3)).astype(F32)   # many particles per cell
    o, _ = oracle_for([{"pos": p, "id": 1000 + i} for i, p in enumerate(pos)], [], L=30.0)
    o.run_stage(sc.STAGE_SORT)
    _, p1, _ = o.snapshot()
    o.run_stage(sc.STAGE_SORT)     # sorting sorted input is the identity
    assert np.array_equal(o.export(sc.EXPORT_PERMUTATION, 0), np.arange(500, dtype=np.int32))
    keys = o.export(sc.EXPORT_KEYS, 0)
    rng_ = o.export(sc.EXPORT_RANGES, 0).reshape(-1, 3)
    brute = []
    for k in np.unique(keys):
        idx = np.nonzero(keys == k)[0]
        assert np.array_equal(idx, np.arange(idx[0], idx[-1] + 1))
        brute.append([k, idx[0], idx[-1] + 1])
    assert np.array_equal(rng_, np.array(brute, dtype=np.uint32))
    # within each cell, particles keep their previous relative order
    ids = p1["id"][:500]
    for k, s, e in brute:
        assert np.all(np.diff(ids[s:e]) > 0)


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_neighbor_lists_brute_force(boundary):
    L, cut = 50.0, 8.0
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, L, (1200, 3)).astype(F32)
    o, _ = oracle_for([{"pos": p} for p in pos], [], L=L, boundary=boundary, cutoff=cut)
    o.run_stage(sc.STAGE_SORT)
    _, p, _ = o.snapshot()
    P = p["pos"][:1200]
    pairs = o.export(sc.EXPORT_NEIGHBORS, 0).reshape(-1, 2)
    got = {}
    for i, j in pairs:
        got.setdefault(int(i), []).append(int(j))
    for i in range(1200):
        d = P - P[i]
        if boundary == "periodic":
            d = np.where(d > L / 2, d - F32(L), np.where(d < -L / 2, d + F32(L), d))
        inside = np.nonzero((np.abs(d) <= cut).all(axis=1))[0]
        inside = inside[inside != i]
        lst = got.get(i, [])
        assert sorted(lst) == sorted(inside.tolist()), i
        if boundary == "clamp":   # cells ascending in key order and slots ascending = ascending slots
            assert lst == sorted(lst)


# ---- integrator of an isolated particle (closed form, float64) -----------------

def u01(x):
    return ((x.astype(np.uint64) >> 8) + 1).astype(np.float64) * 2.0 ** -24


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by + ay * bw + az * bx - ax * bz,
                     aw * bz + az * bw + ax * by - ay * bx, aw * bw - ax * bx - ay * by - az * bz])


def rot_y(q):
    x, y, z, w = q
    return np.array([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)])


def slerp(a, b, t):
    c = float(np.dot(a, b))
    if c < 0:
        b, c = -b, -c
    if c > 1 - 1e-6:
        return (1 - t) * a + t * b
    om = np.arccos(c)
    return (np.sin((1 - t) * om) * a + np.sin(t * om) * b) / np.sin(om)


def shoemake(u, v, w):
    return np.array([np.sqrt(1 - u) * np.sin(2 * np.pi * v), np.sqrt(1 - u) * np.cos(2 * np.pi * v),
                     np.sqrt(u) * np.sin(2 * np.pi * w), np.sqrt(u) * np.cos(2 * np.pi * w)])


@pytest.mark.parametrize("pop,radius", [(0, 2.5), (0, 4.0), (1, 1.0)])
def test_isolated_particle_step_closed_form(pop, radius):
    L = 100.0
    cfgkw = dict(C=3, R=4, D=2)
    q0 = np.array([0.1, -0.3, 0.2, 0.9274], np.float32)
    q0 = (q0 / np.linalg.norm(q0)).astype(np.float32)
    row = {"pos": [40.0, 50.0, 60.0], "radius": radius, "id": 7, "type": 1 if pop == 0 else 1000, "rot": q0,
           "vel": [0.25, -0.5, 0.125]}
    o, s = oracle_for([row] if pop == 0 else [], [row] if pop == 1 else [], L=L, **cfgkw)
    o.step(1)
    _, p, m = o.snapshot()
    got = (p if pop == 0 else m)[0]
    P = s.params
    seed = (P.seed_lo, P.seed_hi)
    r0 = orc.philox([[7, 0, pop, 0]], [seed])[0]
    r1 = orc.philox([[7, 0, pop, 1]], [seed])[0]
    noise = P.movement_noise_scale if pop == 0 else P.metabolite_movement_noise_scale
    d = shoemake(u01(r0[:1])[0], u01(r0[1:2])[0], u01(r0[2:3])[0])
    n = np.sqrt(-2 * np.log(u01(r0[3:4])[0])) * np.cos(2 * np.pi * u01(r1[:1])[0])
    pos_noise = rot_y(d) * abs(n) * np.float64(F32(noise)) / max(radius / 2.5, 1.0)
    rr = shoemake(u01(r1[1:2])[0], u01(r1[2:3])[0], u01(r1[3:4])[0])
    ang_noise = slerp(np.array([0, 0, 0, 1.0]), rr, np.float64(F32(P.rotation_noise_scale)) / (radius / 2.5))
    vel = np.array(row["vel"], np.float64) + pos_noise
    av = ang_noise
    pos = np.clip(np.array(row["pos"]) + vel * 0.45, 0, L)
    rot = qmul(slerp(np.array([0, 0, 0, 1.0]), av, 0.45), q0.astype(np.float64))
    rot /= np.linalg.norm(rot)
    av2 = slerp(av, np.array([0, 0, 0, 1.0]), 1 - 0.7)
    np.testing.assert_allclose(got["pos"], pos, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got["vel"], vel * 0.7, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(got["rot"], rot, atol=2e-6)
    np.testing.assert_allclose(got["ang_vel"], av2, atol=2e-6)
    assert np.array_equal(got["pos_noise"], [0, 0, 0]) and np.array_equal(got["ang_noise"], [0, 0, 0, 1])
    assert abs(np.linalg.norm(got["rot"].astype(np.float64)) - 1) < 1e-6


# ---- two-body closed forms -------------------------------------------------------

def test_collision_pair_relax_closed_form():
    rows = [{"pos": [10.0, 20.0, 20.0], "radius": 3.0, "id": 0}, {"pos": [14.0, 20.0, 20.0], "radius": 3.0, "id": 1}]
    o, _ = oracle_for(rows, [])
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_RELAX)
    _, p, _ = o.snapshot()
    # dVel = (tp.pos - n*(r+r') + tp.vel - p.pos) * 0.9 * r'/(r+r') = (14-6-10)*0.45
    v = {int(p["id"][i]): p["vel"][i] for i in range(2)}
    np.testing.assert_allclose(v[0], [-0.9, 0, 0], atol=1e-6)
    np.testing.assert_allclose(v[1], [0.9, 0, 0], atol=1e-6)


def test_metabolic_lipid_collision_closed_form():
    plain = [{"pos": [20.0, 20.0, 20.0], "radius": 2.5, "type": 0, "hydrophobic": 2, "id": 0}]
    met = [{"pos": [17.0, 20.0, 20.0], "radius": 1.0, "type": 1000, "id": 0}]
    o, _ = oracle_for(plain, met)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_RELAX_METABOLIC)
    o.run_stage(sc.STAGE_RELAX_PARTNERS)
    _, p, m = o.snapshot()
    cd_m = np.float32(3.5 + 0.8 * 1.0)          # r + r' + 0.8 r (metabolic side)
    cd_l = np.float32(3.5 + 0.8 * 2.5)          # lipid side
    # metabolic: weight 2, dVel = (20 - cd_m - 17) * 0.9 * 2, divided by the weight 2
    np.testing.assert_allclose(m["vel"][0], [(3.0 - cd_m) * 0.9, 0, 0], rtol=1e-6)
    # lipid sees the metabolic particle moved: tp.vel now set, weight 2
    mv = m["vel"][0][0]
    np.testing.assert_allclose(p["vel"][0], [(17.0 + cd_l + mv - 20.0) * 0.9, 0, 0], rtol=1e-5)


def test_diffusion_pair_and_dead_band():
    met = [{"pos": [20.0, 20.0, 20.0], "radius": 1.0, "type": 1000, "metabolites": [100, 0.0005, 0, 0, 0, 0, 0, 0, 0, 0]},
           {"pos": [25.0, 20.0, 20.0], "radius": 1.0, "type": 1000, "metabolites": [0, 0, 0, 0, 0, 0, 0, 0, 0, 8]},
           {"pos": [40.0, 40.0, 40.0], "radius": 1.0, "type": 1000, "metabolites": [50] * 10}]
    o, _ = oracle_for([], met)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_DIFFUSE)
    _, _, m = o.snapshot()
    by = {tuple(m["pos"][i]): m["metabolites"][i] for i in range(3)}
    a, b, c = by[(20.0, 20.0, 20.0)], by[(25.0, 20.0, 20.0)], by[(40.0, 40.0, 40.0)]
    assert a[0] == np.float32(90.0) and b[0] == np.float32(10.0)
    assert a[1] == np.float32(0.0005) and b[1] == 0.0       # |diff| <= 0.001: no exchange
    assert a[9] == np.float32(0.8) and b[9] == np.float32(7.2)
    assert np.all(c == 50)


def test_diffusion_conserves_mass_stage_level():
    cfg, setup = scenes.mixed_scene(seed=4)
    s = sc.Setup(cfg, setup)
    o = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic)
    o.run_stage(sc.STAGE_SORT)
    _, _, m0 = o.snapshot()
    for j in range(8):
        o.run_stage(sc.STAGE_DIFFUSE, j)
    _, _, m1 = o.snapshot()
    t0 = m0["metabolites"].astype(np.float64).sum(axis=0)
    t1 = m1["metabolites"].astype(np.float64).sum(axis=0)
    np.testing.assert_allclose(t1, t0, rtol=2e-6, atol=1e-4)
    assert not np.allclose(m0["metabolites"], m1["metabolites"])


# ---- bonds: formation, transitions, polymerisation --------------------------------

def test_bond_formation_transition_and_polymerisation():
    A = {"pos": [20.0, 20.0, 20.0], "radius": 2.0, "type": 100, "id": 0}
    B = {"pos": [20.0, 20.0, 23.0], "radius": 2.0, "type": 101, "id": 1}
    C = {"pos": [20.0, 23.0, 20.0], "radius": 2.0, "type": 101, "id": 2}   # a second B next to A
    o, _ = oracle_for([A, B, C], [], interactions=scenes.REACTIONS)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_COMPLEXIFY, 0)
    _, p, _ = o.snapshot()
    by = {int(p["id"][i]): p[i] for i in range(3)}
    # A (lower type) bonds B first (B has the lower cell key); C is then refused on A's side because
    # A's own group 1001 is already active in this pass ...
    assert by[0]["n_bonds"] == 1 and by[0]["bonds"][0].tolist() == [1001, 1001, 1]
    assert by[1]["n_bonds"] == 1 and by[1]["bonds"][0].tolist() == [1001, 1001, 0]
    # ... but C still bonds A one-sidedly: it sees A's bonds from before the pass (Jacobi), the
    # reference's open TODO at interactions.cuh:298
    assert by[2]["n_bonds"] == 1 and by[2]["bonds"][0].tolist() == [1001, 1001, 0]
    o.run_stage(sc.STAGE_TRANSITION, 5)
    _, p, _ = o.snapshot()
    by = {int(p["id"][i]): p[i] for i in range(3)}
    # A: deferred transition to 1002 sets state 1 and schedules polymerisation 1004
    assert by[0]["bonds"][0].tolist() == [1002, 1001, 1]
    assert by[0]["state"] == 1 and by[0]["last_state_change_step"] == 5
    assert by[0]["scheduled_polymerization"] == 1004
    # B: 1002 leaves B's state alone -> immediate transition, no state change
    assert by[1]["bonds"][0].tolist() == [1002, 1001, 0] and by[1]["state"] == 0
    o.run_stage(sc.STAGE_POLYMERIZE, 5)
    c, p, _ = o.snapshot()
    assert c.n_active == 4 and c.next_particle_id == 4
    new = p[3]
    assert new["id"] == 3 and new["type"] == 102 and new["state"] == 2 and new["last_state_change_step"] == 5
    assert new["n_bonds"] == 1 and new["bonds"][0].tolist() == [1004, 1004, 0]
    a = {int(p["id"][i]): p[i] for i in range(4)}[0]
    assert a["scheduled_polymerization"] == -1
    assert a["n_bonds"] == 2 and a["bonds"][1].tolist() == [1004, 1004, 3]
    # monomer sits at A + R(A) * relativePosition(1004) = (20, 22.5, 20); radius from the type table
    np.testing.assert_allclose(new["pos"], [20.0, 22.5, 20.0], atol=1e-5)
    assert new["radius"] == np.float32(1.5)


def test_break_after_transition_and_capacity():
    rows = [{"pos": [20.0, 20.0, 20.0], "radius": 2.0, "type": 100, "id": 0}]
    o, s = oracle_for(rows + [{"pos": [20.0, 20.0, 23.0], "radius": 2.0, "type": 101, "id": 1}], [],
                      interactions=scenes.REACTIONS)
    c, p, m = o.snapshot()
    p[0]["n_bonds"] = 1
    p[0]["bonds"][0] = [1003, 1001, 1]
    o2 = orc.Oracle(s.params, s.tables, c, p, m)
    o2.run_stage(sc.STAGE_SORT)
    o2.run_stage(sc.STAGE_TRANSITION, 0)
    _, p2, _ = o2.snapshot()
    a = {int(p2["id"][i]): p2[i] for i in range(2)}[0]
    assert a["n_bonds"] == 0       # 1003 breaks after transition


# ---- frames.dat ----------------------------------------------------------------------

def test_frame_layout_matches_reader_offsets():
    cfg, setup = scenes.mixed_scene(seed=8)
    s = sc.Setup(cfg, setup)
    o = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic)
    o.step(1)
    c, p, m = o.snapshot()
    raw = o.emit_frame()
    P = s.params
    assert len(raw) == 4 + P.num_particles * 36 + 4 + P.num_metabolic_particles * 44
    rec_p = np.dtype([("id", "<i4"), ("type", "<i4"), ("flags", "<i4"), ("h", "<f2", (12,))])
    rec_m = np.dtype([("id", "<i4"), ("type", "<i4"), ("flags", "<i4"), ("h", "<f2", (12,)), ("met", "<f2", (4,))])
    assert rec_p.itemsize == 36 and rec_m.itemsize == 44
    n0 = int(np.frombuffer(raw[:4], "<u4")[0])
    fp = np.frombuffer(raw[4:4 + n0 * 36], rec_p)
    off = 4 + n0 * 36
    n1 = int(np.frombuffer(raw[off:off + 4], "<u4")[0])
    fm = np.frombuffer(raw[off + 4:off + 4 + n1 * 44], rec_m)
    assert n0 == P.num_particles and n1 == P.num_metabolic_particles
    na = c.n_active
    assert np.array_equal(fp["id"][:na], p["id"][:na])
    assert np.array_equal(fp["flags"][:na], p["flags"][:na] | ((p["state"][:na] & 0xF) << 16))
    vals = np.concatenate([p["radius"][:na, None], p["pos"][:na], p["rot"][:na], p["debug"][:na]], axis=1)
    assert np.array_equal(fp["h"][:na].view(np.uint16), vals.astype(np.float16).view(np.uint16))
    assert not fp[na:].view(np.uint8).any()            # slots >= nActive are zero records
    nm = c.n_active_metabolic
    assert np.array_equal(fm["met"][:nm].view(np.uint16), m["metabolites"][:nm, :4].astype(np.float16).view(np.uint16))


# ---- polar alignment (dynamics.cuh:478-582): invariants ----------------------------

def _qrot(q, v):
    """Rotate v by unit quaternion q = (x, y, z, w), float64."""
    x, y, z, w = (float(c) for c in q)
    u = np.array([x, y, z])
    v = np.asarray(v, dtype=np.float64)
    return v + 2.0 * w * np.cross(u, v) + 2.0 * np.cross(u, np.cross(u, v))


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + bw * ax + ay * bz - az * by, aw * by + bw * ay + az * bx - ax * bz,
                     aw * bz + bw * az + ax * by - ay * bx, aw * bw - ax * bx - ay * by - az * bz])


def _angle(a, b):
    a = np.asarray(a, np.float64) / np.linalg.norm(a)
    b = np.asarray(b, np.float64) / np.linalg.norm(b)
    return float(np.arccos(np.clip(np.dot(a, b), -1.0, 1.0)))


def _polar_pair(d, q):
    """Two polar lipids (r 2.5) at distance d along x about the box centre,
    orientations q and Rz(pi) q: the pair is symmetric under a half turn about z."""
    rz = np.array([0.0, 0.0, 1.0, 0.0])
    qb = _qmul(rz, q)
    c = 30.0
    rows = [{"pos": [c - d / 2, c, c], "radius": 2.5, "type": 0, "hydrophobic": 2, "id": 0, "rot": list(q)},
            {"pos": [c + d / 2, c, c], "radius": 2.5, "type": 0, "hydrophobic": 2, "id": 1, "rot": list(qb)}]
    o, _ = oracle_for(rows, [])
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_RELAX)
    _, p, _ = o.snapshot()
    by_id = {int(p["id"][k]): k for k in range(2)}
    return rows, p[by_id[0]], p[by_id[1]]


def test_polar_alignment_symmetry_threshold_and_realignment():
    """Pins the hydrophobic-alignment block without restating it: (1) a pair
    symmetric under a half turn about z gets velocity and angular-velocity
    updates related by that half turn (a wrong sign, a swapped operand or the
    partner's instead of the own angular velocity breaks it); (2) the block
    switches on at dist <= 1.4 (r_p + r_t) = 7 and below the collision distance
    5 nothing else acts; (3) the realignment turns the particle's -up axis 20 %
    of the way towards the direction of the partner's displaced polar end."""
    rng = np.random.default_rng(11)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    rows, a, b = _polar_pair(6.0, q)
    va, vb = a["vel"].astype(np.float64), b["vel"].astype(np.float64)
    assert np.linalg.norm(va) > 1e-4                                  # the alignment term acted
    np.testing.assert_allclose(vb, [-va[0], -va[1], va[2]], rtol=1e-5, atol=1e-6)
    wa, wb = a["ang_vel"].astype(np.float64), b["ang_vel"].astype(np.float64)
    assert abs(wa[3]) < 1.0 - 1e-7                                    # and turned the particle
    np.testing.assert_allclose(wb, [-wa[0], -wa[1], wa[2], wa[3]], rtol=1e-5, atol=1e-6)

    # threshold: just outside 1.4 (r_p + r_t) nothing changes (no collision at > 5 either)
    _, a2, b2 = _polar_pair(7.05, q)
    assert not a2["vel"].any() and not b2["vel"].any()
    np.testing.assert_array_equal(a2["ang_vel"], [0, 0, 0, 1])
    _, a3, _ = _polar_pair(6.95, q)
    assert np.linalg.norm(a3["vel"].astype(np.float64)) > 0

    # realignment: -up of A is turned 20 % of the way towards the partner's polar end
    # (partner at rest, own angular velocity identity: tPDV = t.pos - r_t * up_t)
    up_a = _qrot(rows[0]["rot"], [0, 1, 0])
    up_b = _qrot(rows[1]["rot"], [0, 1, 0])
    target = np.array(rows[1]["pos"]) - 2.5 * up_b
    direction = target - (np.array(rows[0]["pos"]) + va)
    before = _angle(-up_a, direction)
    after = _angle(_qrot(wa, -up_a), direction)
    assert before > 0.05
    np.testing.assert_allclose(after, 0.8 * before, rtol=1e-4)


def test_alignment_uses_own_angular_velocity_only():
    """dynamics.cuh:476 rotates the partner's up vector by p's own angular
    velocity: p's update must not change when only the partner's angular
    velocity changes (and must change when p's does)."""
    rng = np.random.default_rng(5)
    qa, qb = (v / np.linalg.norm(v) for v in rng.normal(size=(2, 4)))

    def run(wa, wb):
        rows = [{"pos": [27.0, 30.0, 30.0], "radius": 2.5, "type": 0, "hydrophobic": 2, "id": 0, "rot": list(qa)},
                {"pos": [33.0, 30.5, 30.0], "radius": 2.5, "type": 0, "hydrophobic": 2, "id": 1, "rot": list(qb)}]
        o, s = oracle_for(rows, [])
        c, p, m = o.snapshot()
        p["ang_vel"][0], p["ang_vel"][1] = wa, wb
        o2 = orc.Oracle(s.params, s.tables, c, p, m)
        o2.run_stage(sc.STAGE_SORT)
        o2.run_stage(sc.STAGE_RELAX)
        _, q, _ = o2.snapshot()
        k = [int(x) for x in q["id"][:2]].index(0)
        return q[k]["vel"].tobytes() + q[k]["ang_vel"].tobytes()

    turn = lambda t: [0.0, np.sin(t / 2), 0.0, np.cos(t / 2)]
    base = run([0, 0, 0, 1], [0, 0, 0, 1])
    assert run([0, 0, 0, 1], turn(0.3)) == base
    assert run(turn(0.3), [0, 0, 0, 1]) != base


# ---- propagateState (interactions.cuh:743-822) ----------------------------------

def _propagate(pos, states, bonds, interactions, cutoff=8.0):
    rows = [{"pos": p, "id": 10 + k, "state": st, "last_state_change_step": lc, "bonds": bonds[k]}
            for k, (p, (st, lc)) in enumerate(zip(pos, states))]
    o, _ = oracle_for(rows, [], interactions=interactions, cutoff=cutoff)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_PROPAGATE_STATE)
    _, p, _ = o.snapshot()
    got = {int(p["id"][k]): (int(p["state"][k]), int(p["last_state_change_step"][k])) for k in range(len(rows))}
    return [got[10 + k] for k in range(len(rows))]


def _links(ids, bond_type):
    """Serial links of a chain, as the reference's linkParticlesSerially makes them."""
    out = [[] for _ in ids]
    for k in range(1, len(ids)):
        out[k - 1].append([bond_type, 10, ids[k]])
        out[k].append([bond_type, 11, ids[k - 1]])
    return out


def test_propagate_state_takes_the_latest_partner_state():
    """A chain A-B-C, 5 nm apart: each particle takes the state of a linked
    partner that changed state later than itself, comparing against the
    state read at the start of the pass (A gets B's, B gets C's, C keeps
    its own); links whose interaction does not propagate state, chain
    links (type < 10) and partners beyond the interaction distance change
    nothing."""
    rows = [{"id": 1001, "group": 1001, "firstPartnerType": 1, "secondPartnerType": 1, "propagateState": True,
             "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 5]},
            {"id": 1002, "group": 1002, "firstPartnerType": 1, "secondPartnerType": 1,
             "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 5]}]
    pos = [[30.0, 30.0, 25.0], [30.0, 30.0, 30.0], [30.0, 30.0, 35.0]]
    states = [(0, -1), (3, 7), (5, 9)]
    ids = [10, 11, 12]
    assert _propagate(pos, states, _links(ids, 1001), rows) == [(3, 7), (5, 9), (5, 9)]
    assert _propagate(pos, states, _links(ids, 1002), rows) == states        # no propagateState flag
    assert _propagate(pos, states, _links(ids, 1), rows) == states           # chain link, type < 10
    far = [[30.0, 30.0, 20.0], [30.0, 30.0, 30.0], [30.0, 30.0, 40.0]]     # 10 nm apart, cutoff 8
    assert _propagate(far, states, _links(ids, 1001), rows) == states
    # B's partners A (cell below) and C both changed at step 9: B takes the
    # first one visited in stencil order, and C's equal step does not replace
    # it; A and C see no later partner and keep their own
    assert _propagate(pos, [(1, 9), (3, 7), (5, 9)], _links(ids, 1001), rows) == [(1, 9), (1, 9), (5, 9)]


# ---- bonded partners: special cases with closed forms -------------------------------
# relax between bonded partners (dynamics.cuh:585-770) and coordinateNoise
# (dynamics.cuh:288-402).  Expected values are worked out by hand from the
# geometry of each case (targets that coincide with the current pose, tilts
# inside and beyond the chain's pi/8 flex), not from the code's formulas.

def _rot(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([axis * np.sin(ang / 2), [np.cos(ang / 2)]])


def _same_rotation(q, e, atol=2e-6):
    q = np.asarray(q, np.float64)
    assert np.allclose(q, e, atol=atol) or np.allclose(q, -np.asarray(e), atol=atol), (q, e)


ALPHA = 0.6
COMPLEX_RX = [
    {"id": 1030, "group": 1030, "firstPartnerType": 100, "secondPartnerType": 101,
     "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 6.0, 0]},
    {"id": 1031, "group": 1031, "firstPartnerType": 100, "secondPartnerType": 101,
     "relativeOrientation": _rot([0, 1, 0], ALPHA).tolist(), "relativePosition": [0, 0, 0]},
]


def _relax_once(rows, interactions):
    o, _ = oracle_for(rows, [], interactions=interactions)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_RELAX, 0)
    _, p, _ = o.snapshot()
    return {int(p["id"][i]): p[i] for i in range(len(rows))}


@pytest.mark.parametrize("beta", [0.0, 0.7])
def test_complex_bond_position_target(beta):
    """p (type 100, r 2) sits delta away from where its partner t (type 101,
    r 3) wants it: t.pos - R(t) (0, 6, 0), R(t) the common orientation of both.
    Each side's pull is the mirrored average of the two targets: p moves by
    -0.405 r_p/(r_p+r_t) delta, t by +0.405 r_t/(r_p+r_t) delta."""
    R = _rot([1, 0, 0], beta)
    T = np.array([30.0, 30.0, 30.0])
    delta = np.array([0.3, -0.2, 0.4])
    up6 = qrot(R, np.array([0.0, 6.0, 0.0]))
    P = T - up6 + qrot(R, delta)
    rows = [{"pos": P.tolist(), "radius": 2.0, "type": 100, "id": 0, "rot": R.tolist(), "bonds": [(1030, 1030, 1)]},
            {"pos": T.tolist(), "radius": 3.0, "type": 101, "id": 1, "rot": R.tolist(), "bonds": [(1030, 1030, 0)]}]
    by = _relax_once(rows, COMPLEX_RX)
    np.testing.assert_allclose(by[0]["vel"], -0.405 * 0.4 * qrot(R, delta), atol=2e-6)
    np.testing.assert_allclose(by[1]["vel"], 0.405 * 0.6 * qrot(R, delta), atol=2e-6)
    for k in (0, 1):   # orientations already match (identity relative orientation)
        _same_rotation(by[k]["ang_vel"], [0, 0, 0, 1])


@pytest.mark.parametrize("beta", [0.0, 0.9])
def test_complex_bond_orientation_target(beta):
    """Relative orientation rot_y(alpha) between two partners of the same
    orientation: p turns by -0.9 r_t/(r_p+r_t) alpha about y, t by
    +0.9 r_p/(r_p+r_t) alpha; a common orientation rot_x(beta) of both
    changes nothing (the targets are relative)."""
    R = _rot([1, 0, 0], beta)
    rows = [{"pos": [30.0, 30.0, 30.0], "radius": 2.0, "type": 100, "id": 0, "rot": R.tolist(),
             "bonds": [(1031, 1031, 1)]},
            {"pos": [30.0, 30.0, 34.0], "radius": 3.0, "type": 101, "id": 1, "rot": R.tolist(),
             "bonds": [(1031, 1031, 0)]}]
    by = _relax_once(rows, COMPLEX_RX)
    _same_rotation(by[0]["ang_vel"], _rot([0, 1, 0], -0.9 * 0.6 * ALPHA))
    _same_rotation(by[1]["ang_vel"], _rot([0, 1, 0], 0.9 * 0.4 * ALPHA))


def test_complex_bond_angular_velocities_meet():
    """Same orientation, identity relative orientation; p does not spin, t
    spins rot_y(theta): each is turned towards the other's spin by 0.9 times
    the partner's radius share: p to rot_y(0.54 theta), t to rot_y(0.64
    theta) (= rot_y(-0.36 theta) after its own rot_y(theta))."""
    theta = 0.5
    rows = [{"pos": [30.0, 30.0, 30.0], "radius": 2.0, "type": 100, "id": 0, "bonds": [(1031, 1031, 1)]},
            {"pos": [30.0, 30.0, 34.0], "radius": 3.0, "type": 101, "id": 1, "bonds": [(1031, 1031, 0)],
             "ang_vel": _rot([0, 1, 0], theta).tolist()}]
    rx = [dict(COMPLEX_RX[1], relativeOrientation=[0, 0, 0, 1])]
    by = _relax_once(rows, rx)
    _same_rotation(by[0]["ang_vel"], _rot([0, 1, 0], 0.54 * theta))
    _same_rotation(by[1]["ang_vel"], _rot([0, 1, 0], 0.64 * theta))


def _qpow(q, t):
    """q^t for a unit quaternion (the short arc): the same axis, t times the angle."""
    q = np.asarray(q, np.float64)
    if q[3] < 0:
        q = -q
    ang = 2 * np.arccos(min(q[3], 1.0))
    if ang < 1e-12:
        return np.array([0, 0, 0, 1.0])
    return _rot(q[:3], t * ang)


def test_complex_bond_orientation_midpoint():
    """Relative orientation rot_y(alpha), partner t spinning rot_x(a), p still:
    p's view of the error is rel = rot_x(a) rot_y(-alpha), t's is other =
    rot_x(-a) rot_y(alpha).  Their slerp midpoint is (rel + other)/|rel + other|
    = rot_z(-2 atan(tan(a/2) tan(alpha/2))), and p turns 0.54 of the way along
    midpoint * rel (the midpoint applied after rel)."""
    a, alpha = 0.6, 0.8
    rows = [{"pos": [30.0, 30.0, 30.0], "radius": 2.0, "type": 100, "id": 0, "bonds": [(1031, 1031, 1)]},
            {"pos": [30.0, 30.0, 34.0], "radius": 3.0, "type": 101, "id": 1, "bonds": [(1031, 1031, 0)],
             "ang_vel": _rot([1, 0, 0], a).tolist()}]
    rx = [dict(COMPLEX_RX[1], relativeOrientation=_rot([0, 1, 0], alpha).tolist())]
    by = _relax_once(rows, rx)
    rel = qmul(_rot([1, 0, 0], a), _rot([0, 1, 0], -alpha))
    mid = _rot([0, 0, 1], -2 * np.arctan(np.tan(a / 2) * np.tan(alpha / 2)))
    _same_rotation(by[0]["ang_vel"], _qpow(qmul(mid, rel), 0.54), atol=3e-6)


def test_type0_bond_turns_rotation_only():
    """A bond of type 0 (the reference's test interaction) turns p's own
    rotation 20 % of the way to its fixed relative orientation to the partner,
    trod = rot_y(-pi/6) rot_x(pi/6) (p, of the lower type, towards inv(trod);
    t towards trod), and adds no velocity."""
    rows = [{"pos": [30.0, 30.0, 30.0], "radius": 2.0, "type": 100, "id": 0, "bonds": [(0, 0, 1)]},
            {"pos": [30.0, 30.0, 34.0], "radius": 3.0, "type": 101, "id": 1, "bonds": [(0, 0, 0)]}]
    by = _relax_once(rows, [])
    trod = qmul(_rot([0, 1, 0], -np.pi / 6), _rot([1, 0, 0], np.pi / 6))
    inv = trod * np.array([-1, -1, -1, 1.0])
    _same_rotation(by[0]["rot"], _qpow(inv, 0.2))
    _same_rotation(by[1]["rot"], _qpow(trod, 0.2))
    np.testing.assert_allclose(by[0]["vel"], [0, 0, 0], atol=0)
    np.testing.assert_allclose(by[1]["vel"], [0, 0, 0], atol=0)


def qrot(q, v):
    return _qrot(np.asarray(q, np.float64), np.asarray(v, np.float64))


def _chain_pair(t_rel):
    """p (r 2, up +y, FORWARD link) at (30, 30, 30) and t (r 3, up +y,
    BACKWARD link) at p + t_rel: a chain wants t straight below p (along
    -up(p)) and p straight above t, each within pi/8."""
    P = np.array([30.0, 30.0, 30.0])
    rows = [{"pos": P.tolist(), "radius": 2.0, "type": 100, "id": 0, "bonds": [(1, 10, 1)]},
            {"pos": (P + t_rel).tolist(), "radius": 3.0, "type": 101, "id": 1, "bonds": [(1, 11, 0)]}]
    return _relax_once(rows, [])


def test_chain_link_radial_stretch():
    """Aligned link stretched from 5 (= r_p + r_t) to 5.5: both ends are pulled
    together along the link by 0.405 (5.5 - 5) times their own radius share;
    no turn; p's debug vector records 0.1 (target - position)."""
    by = _chain_pair(np.array([0.0, -5.5, 0.0]))
    np.testing.assert_allclose(by[0]["vel"], [0, -0.405 * 0.5 * 0.4, 0], atol=2e-6)
    np.testing.assert_allclose(by[1]["vel"], [0, 0.405 * 0.5 * 0.6, 0], atol=2e-6)
    _same_rotation(by[0]["ang_vel"], [0, 0, 0, 1])
    _same_rotation(by[1]["ang_vel"], [0, 0, 0, 1])
    np.testing.assert_allclose(by[0]["debug"][:3], [0, -0.05, 0], atol=2e-6)


def test_chain_link_tilt_within_flex_is_free():
    """A link of the right length tilted by pi/16 (< pi/8) from the up axes:
    the targets are the current positions, so no pull; p still turns its up
    axis towards the link by 0.9 r_t/(r_p+r_t) of the tilt (about -z)."""
    phi = np.pi / 16
    u = np.array([np.sin(phi), np.cos(phi), 0.0])   # from t to p
    by = _chain_pair(-5.0 * u)
    np.testing.assert_allclose(by[0]["vel"], [0, 0, 0], atol=2e-6)
    np.testing.assert_allclose(by[1]["vel"], [0, 0, 0], atol=2e-6)
    _same_rotation(by[0]["ang_vel"], _rot([0, 0, -1], 0.9 * 0.6 * phi))


def test_chain_link_tilt_beyond_flex():
    """Tilted by pi/4: each side's target is its partner's up axis turned
    towards the link by pi/8 only, so p is pulled by 0.405 r_p/(r_p+r_t) times
    5 ((sin, cos)(pi/8) - (sin, cos)(pi/4)) in (x, y)."""
    phi = np.pi / 4
    u = np.array([np.sin(phi), np.cos(phi), 0.0])
    by = _chain_pair(-5.0 * u)
    v = 5.0 * (np.array([np.sin(np.pi / 8), np.cos(np.pi / 8), 0.0]) - u)
    np.testing.assert_allclose(by[0]["vel"], 0.405 * 0.4 * v, atol=3e-6)
    np.testing.assert_allclose(by[1]["vel"], -0.405 * 0.6 * v, atol=3e-6)
    _same_rotation(by[0]["ang_vel"], _rot([0, 0, -1], 0.9 * 0.6 * phi))


def _coordinate_once(rows, arg=0, C=16):
    o, _ = oracle_for(rows, [], C=C)
    o.run_stage(sc.STAGE_SORT)
    o.run_stage(sc.STAGE_COORDINATE, arg)
    _, p, _ = o.snapshot()
    return {int(p["id"][i]): p[i] for i in range(len(rows))}


def test_coordinate_noise_pair():
    """p (r 2) noise (1, 0, 0), partner t (r 3) noise (0, 1, 0): the pull
    0.54 (t - p) = (-0.54, 0.54, 0) is shorter than |p's noise| = 1, so it is
    stretched to length 1: p's noise becomes (1 - 1/sqrt2, 1/sqrt2, 0).  t's
    pull 0.36 (p - t) is stretched to |t's noise| = 1 as well.  Angular: p's
    noise turns by 0.54 of the partner's relative turn, composed on p's side
    (p = rot_x(0.5), t = rot_x(0.5) rot_z(0.8): p -> rot_z(0.432) rot_x(0.5))."""
    qa = _rot([1, 0, 0], 0.5)
    qb = qmul(qa, _rot([0, 0, 1], 0.8))
    rows = [{"pos": [30.0, 30.0, 30.0], "radius": 2.0, "type": 100, "id": 0, "bonds": [(1030, 1030, 1)],
             "pos_noise": [1.0, 0.0, 0.0], "ang_noise": qa.tolist()},
            {"pos": [30.0, 33.0, 30.0], "radius": 3.0, "type": 101, "id": 1, "bonds": [(1030, 1030, 0)],
             "pos_noise": [0.0, 1.0, 0.0], "ang_noise": qb.tolist()}]
    by = _coordinate_once(rows)
    s2 = np.sqrt(0.5)
    np.testing.assert_allclose(by[0]["pos_noise"], [1 - s2, s2, 0], atol=2e-6)
    np.testing.assert_allclose(by[1]["pos_noise"], [s2, 1 - s2, 0], atol=2e-6)
    _same_rotation(by[0]["ang_noise"], qmul(_rot([0, 0, 1], 0.54 * 0.8), qa))


def test_coordinate_noise_long_pull_and_chain_fraction():
    """A pull longer than p's own noise is kept as it is (p noise (0.1, 0, 0),
    t (0, 2, 0): p -> 0.1 + 0.54 (-0.1, 2, 0)).  A chain link (type 1) takes
    part only while the coordination fraction j / C is at most 0.25; a
    particle in range that is not a partner never does."""
    base = [{"pos": [30.0, 30.0, 30.0], "radius": 2.0, "type": 100, "id": 0, "pos_noise": [0.1, 0.0, 0.0]},
            {"pos": [30.0, 33.0, 30.0], "radius": 3.0, "type": 101, "id": 1, "pos_noise": [0.0, 2.0, 0.0]},
            {"pos": [33.0, 30.0, 30.0], "radius": 3.0, "type": 101, "id": 2, "pos_noise": [0.0, 0.0, 5.0]}]
    rows = [dict(base[0], bonds=[(1030, 1030, 1)]), dict(base[1], bonds=[(1030, 1030, 0)]), base[2]]
    by = _coordinate_once(rows)
    np.testing.assert_allclose(by[0]["pos_noise"], [0.1 - 0.054, 1.08, 0], atol=2e-6)
    np.testing.assert_allclose(by[2]["pos_noise"], [0, 0, 5.0], atol=0)
    rows = [dict(base[0], bonds=[(1, 10, 1)]), dict(base[1], bonds=[(1, 11, 0)]), base[2]]
    by = _coordinate_once(rows, arg=25, C=100)            # 0.25: still coordinated
    np.testing.assert_allclose(by[0]["pos_noise"], [0.1 - 0.054, 1.08, 0], atol=2e-6)
    by = _coordinate_once(rows, arg=26, C=100)            # 0.26 > 0.25: chain links skip
    np.testing.assert_allclose(by[0]["pos_noise"], [0.1, 0, 0], atol=0)
    np.testing.assert_allclose(by[1]["pos_noise"], [0, 2.0, 0], atol=0)


# ---- the OpenMP build of the oracle ------------------------------------------------

def test_threaded_oracle_equals_serial():
    """The pure Jacobi passes over OpenMP threads (liboracle_omp.so) give the
    serial build's state bit for bit, reactions and bonds included."""
    cfg, setup = scenes.mixed_scene(scale=1.0, L=60.0, boundary="periodic", seed=21, init=True)
    s = sc.Setup(cfg, setup)
    states = []
    for threaded in (False, True):
        o = orc.Oracle(s.params, s.tables, s.counts, s.plain, s.metabolic, threaded=threaded)
        o.step(4)
        states.append(o.snapshot())
        o.close()
    (ca, pa, ma), (cb, pb, mb) = states
    scenes.assert_same_state("serial-vs-threaded", ca, pa, ma, cb, pb, mb)
