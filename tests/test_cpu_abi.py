"""C ABI and host logic without a GPU: the library loads and exports every
function include/spatial_cell.h declares; scene construction, config
validation and table flattening follow the reference loaders
(config.cuh:46-75, chemistry.cuh:10-51, interactions.cuh:17-145)."""
import json
import os
import re

import numpy as np
import pytest

import scenes
import spatial_cell as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "spatial_cell.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sc.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    for n in sc._SIGNATURES:
        assert n in names, f"binding uses undeclared {n}"


def test_struct_layouts():
    import ctypes
    assert ctypes.sizeof(sc.Params) == 96
    assert ctypes.sizeof(sc.Counts) == 16
    assert sc.PARTICLE_DTYPE.itemsize == 272
    assert sc.INTERACTION_DTYPE.itemsize == 100
    assert ctypes.sizeof(sc.RequestInfo) == 64


def test_setup_counts_ids_and_generators():
    cfg, setup = scenes.mixed_scene(seed=1)
    s = sc.Setup(cfg, setup)
    P = s.params
    assert s.counts.n_active == len(s.plain) and s.counts.n_active_metabolic == len(s.metabolic)
    assert s.counts.next_particle_id == len(s.plain) and s.counts.step == 0
    assert np.array_equal(s.plain["id"], np.arange(len(s.plain)))
    assert np.array_equal(s.metabolic["id"], np.arange(len(s.metabolic)))
    assert (s.plain["flags"] == 1).all() and (s.plain["scheduled_polymerization"] == -1).all()
    assert (s.plain["pos"] >= 0).all() and (s.plain["pos"] <= P.sim_size).all()
    q = s.plain["rot"].astype(np.float64)
    assert np.abs(np.linalg.norm(q, axis=1) - 1).max() < 1e-6
    # radius / hydrophobic from the type table (chemistry.cuh:33-51)
    lip = s.plain[s.plain["type"] == 0]
    assert (lip["radius"] == np.float32(2.5)).all() and (lip["hydrophobic"] == 2).all()
    # chains linked serially (setup/interaction_setup.cuh:13-43): inner members have 2 bonds
    ch = s.plain[(s.plain["type"] == 1) & (s.plain["n_bonds"] > 0)]
    assert set(np.unique(ch["n_bonds"])) == {1, 2}
    first = ch[ch["bonds"][:, 0, 1] == 10]
    assert (first["bonds"][:, 0, 2] == first["id"] + 1).all()
    # metabolite seeding: value where x < 0.2 L, slot 0 skipped (metabolite_setup.cuh:22)
    m = s.metabolic
    sel = m["pos"][1:, 0] < np.float32(0.2) * P.sim_size
    assert (m["metabolites"][1:, 0][sel] == 100).all() and (m["metabolites"][1:, 0][~sel] == 0).all()
    assert m["metabolites"][0, 0] == 0


def test_same_seed_same_scene_other_seed_differs():
    a = sc.Setup(*scenes.mixed_scene(seed=5))
    b = sc.Setup(*scenes.mixed_scene(seed=5))
    c = sc.Setup(*scenes.mixed_scene(seed=6))
    assert a.plain.tobytes() == b.plain.tobytes()
    assert a.plain.tobytes() != c.plain.tobytes()


def test_grid_and_config_validation():
    ok = scenes.config(10, 10, 400.0, boundary="periodic", cutoff=10.0, diff=10.0)
    sc.Setup(ok, {"types": scenes.TYPES})
    bad = dict(ok, simSize=405.0)
    with pytest.raises(sc.SpatialCellError) as e:
        sc.Setup(bad, {"types": scenes.TYPES})
    assert e.value.status == 1 and "periodic" in str(e.value)
    with pytest.raises(sc.SpatialCellError):
        sc.Setup(dict(ok, boundary="torus"), {})
    with pytest.raises(sc.SpatialCellError):
        sc.Setup(dict(ok, gridCellSize=0.0), {})
    with pytest.raises(sc.SpatialCellError):
        sc.Setup(dict(ok, simSize=20000.0, boundary="clamp"), {})    # > 1024 cells per axis
    with pytest.raises(sc.SpatialCellError) as e:
        sc.Setup(scenes.config(5, 5, 100.0), {"types": scenes.TYPES,
                                              "populations": [{"array": "plain", "type": 0, "count": 6}]})
    assert "capacity" in str(e.value)
    with pytest.raises(sc.SpatialCellError):
        sc.Setup("{not json", {})
    with pytest.raises(sc.SpatialCellError):
        sc.Setup(scenes.config(5, 5, 100.0), {"populations": [{"array": "plain", "type": 77, "count": 1}]})


def test_interaction_loader_defaults_nulls_and_partner_map():
    rows = [
        {"id": 1002, "group": 7, "firstPartnerType": 100, "secondPartnerType": 101,
         "relativeOrientation": [0, 0, 0, 1], "relativePosition": [1, 2, 3], "polymerize": None,
         "transitionTo": None},
        {"id": 1001, "firstPartnerType": 101, "secondPartnerType": 102, "firstPartnerState": 300,
         "relativeOrientation": [0, 0, 0, 1], "relativePosition": [0, 0, 0]},
    ]
    s = sc.Setup(scenes.config(4, 4, 100.0), {"types": scenes.TYPES, "interactions": rows})
    flat, part, pidx = s.tables["flat"], s.tables["partner"], s.tables["partner_index"]
    assert len(flat) == 1003
    r = flat[1002]
    # present-but-null keys read as 0 (jsoncpp asInt of null), absent keys take the default
    assert r["polymerize"] == 0 and r["transition_to"] == 0
    r1 = flat[1001]
    assert r1["group"] == -1 and r1["transition_to"] == -1 and r1["polymerize"] == -1
    assert r1["set_state"] == 1 and r1["wait_for_state"] == 1 and r1["wait_for_alignment"] == 1
    assert r1["only_via_transition"] == 0 and r1["break_after_transition"] == 0
    assert np.isclose(r1["orientation_alignment_threshold"], 0.8) and np.isclose(r1["position_alignment_threshold"], 0.8)
    assert r1["first_partner_state"] == 300 - 256     # stored as int8, like the reference field
    assert r1["second_partner_state"] == -1
    assert flat[5]["id"] == 0                         # unused ids are zero rows
    # partner map: per type ascending, rows by ascending id, each row once per partner type
    s101, e101 = pidx[2 * 101], pidx[2 * 101 + 1]
    assert [x["id"] for x in part[s101:e101]] == [1001, 1002]
    s100, e100 = pidx[2 * 100], pidx[2 * 100 + 1]
    assert [x["id"] for x in part[s100:e100]] == [1002]
    assert pidx[2 * 0] == pidx[2 * 0 + 1]             # no rows for lipids


def test_reference_json_files_parse(tmp_path):
    """The loader reads documents in the reference's gen/*.json format from a
    path (types_file / interactions_file)."""
    types_path = tmp_path / "types.json"
    inter_path = tmp_path / "inter.json"
    types_path.write_text(json.dumps([{"type": 0, "category": "lipid", "name": "Generic lipid", "radius": 1.0,
                                       "hydrophobic": 2},
                                      {"type": 1, "category": "dummy", "name": "Generic molecule", "radius": 2.5}]))
    inter_path.write_text(json.dumps(scenes.REACTIONS))
    s = sc.Setup(scenes.config(100, 10, 100.0), {"types_file": str(types_path), "interactions_file": str(inter_path),
                                                  "populations": [{"array": "plain", "type": 1, "count": 3}]})
    assert s.tables["types"][0]["hydrophobic"] == 2 and s.tables["types"][1]["radius"] == np.float32(2.5)
    assert len(s.tables["flat"]) == max(r["id"] for r in scenes.REACTIONS) + 1
    with pytest.raises(sc.SpatialCellError):
        sc.Setup(scenes.config(10, 10, 100.0), {"types_file": str(tmp_path / "missing.json")})


def test_c4_workload_matches_fixture():
    """The batched workload's instance sizes are a committed fixture
    (tests/golden/c4_sizes.json, made by make_c4_sizes.py)."""
    with open(os.path.join(ROOT, "tests", "golden", "c4_sizes.json")) as f:
        fx = json.load(f)
    sizes = scenes.c4_sizes()
    assert sizes == fx["sizes"] and sum(sizes) == 1 << 24 and len(sizes) == 256
    assert min(sizes) >= 64
    cfg, setup = scenes.c1_mix_scene(sizes[0], seed=42)
    assert cfg["numParticles"] + cfg["numMetabolicParticles"] == sizes[0]
    assert cfg["simSize"] % 10.0 == 0 and cfg["simSize"] >= 30.0
