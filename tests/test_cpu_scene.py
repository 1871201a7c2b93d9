"""The reference scene's placements (setup/setup.cuh with
setup/particle_setup.cuh and setup/interaction_setup.cuh) as generated by
the library's host scene builder: counts, sphere geometry and orientation,
chain spacing and serial links, lines, complexes and metabolite seeding."""
import numpy as np
import pytest

import scenes
import spatial_cell as sc

CHAIN = [161, 160, 145, 160, 145, 160, 145, 164]


def rotate(v, q):
    """Rotation of v by unit quaternion q = (x, y, z, w), rows of q."""
    q = np.asarray(q, dtype=np.float64)
    u, w = q[..., :3], q[..., 3:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


@pytest.fixture(scope="module")
def scene():
    cfg, doc = scenes.reference_scene()
    s = sc.Setup(cfg, doc)
    radius = {t["type"]: t.get("radius", 0.0) for t in doc["types"]}
    complex1 = [c for c in doc["complexes"] if c["id"] == 1][0]["participants"]
    return s, radius, complex1


def layout(complex1):
    n_dna = int(65536 * 0.1)
    sizes = [n_dna, 1936, 1600, len(CHAIN), len(complex1), len(CHAIN), len(complex1), 1000 * len(CHAIN)]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    return sizes, starts


def test_counts_and_ids(scene):
    s, radius, complex1 = scene
    sizes, starts = layout(complex1)
    assert len(s.plain) == starts[-1] == 6553 + 1936 + 1600 + 8 + 89 + 8 + 89 + 8000
    assert len(s.metabolic) == int(65536 * 0.95)
    assert np.array_equal(s.plain["id"], np.arange(len(s.plain)))
    assert (s.plain["type"][:sizes[0]] == 1).all() and (s.metabolic["type"] == 1000).all()
    assert np.allclose(s.plain["radius"][:sizes[0]], radius[1])
    P = s.params
    assert (s.plain["pos"] >= 0).all() and (s.plain["pos"] <= P.sim_size).all()


@pytest.mark.parametrize("layer,outward", [(1, True), (2, False)])
def test_lipid_half_layers(scene, layer, outward):
    """Points on a sphere about (0.5, 0.5, 0.7) L of radius
    sqrt(n / 1000 (r / 2.5)^2) 40; +y turned to the outward normal (outer
    layer) or, after the half turn about x, to the inward one (inner layer)."""
    s, radius, complex1 = scene
    sizes, starts = layout(complex1)
    n = sizes[layer]
    p = s.plain[starts[layer]:starts[layer + 1]]
    centre = np.array([200.0, 200.0, 280.0])
    r = np.sqrt(n / 1000.0 * (radius[0] / 2.5) ** 2) * 40.0
    d = p["pos"].astype(np.float64) - centre
    dist = np.linalg.norm(d, axis=1)
    assert np.abs(dist - r).max() < 1e-4 * r
    # Fibonacci spiral: heights evenly spaced in (-1, 1)
    z = d[:, 2] / dist
    assert np.allclose(np.sort(z), -1.0 + 2.0 * np.arange(1, n + 1) / (n + 1), atol=1e-5)
    up = rotate(np.array([0.0, 1.0, 0.0]), p["rot"])
    cosang = np.sum(up * d / dist[:, None], axis=1)
    assert (cosang > 0.9999).all() if outward else (cosang < -0.9999).all()
    assert (p["type"] == 0).all() and (p["hydrophobic"] == 2).all()


def _check_links(p, bond_type=1):
    n = len(p)
    assert list(p["n_bonds"]) == [1] + [2] * (n - 2) + [1]
    for k in range(n - 1):          # forward link to the next member, backward link from it
        fwd = p["bonds"][k][1 if k > 0 else 0]
        assert list(fwd) == [bond_type, 10, p["id"][k + 1]]
        assert list(p["bonds"][k + 1][0]) == [bond_type, 11, p["id"][k]]


def test_wrapped_chains(scene):
    """Members r(type - 1) + r(type) apart (setup/particle_setup.cuh:82-117),
    each along its own turned up vector, linked serially."""
    s, radius, complex1 = scene
    sizes, starts = layout(complex1)
    chains = [s.plain[starts[3]:starts[4]]]
    big = s.plain[starts[7]:starts[8]]
    chains += [big[k * 8:(k + 1) * 8] for k in range(1000)]
    first = chains[0]["pos"][0].astype(np.float64)
    assert np.allclose(first, [80.0, 200.0, 200.0])
    starts_random = np.array([c["pos"][0] for c in chains[1:]])
    assert starts_random.std(axis=0).min() > 80.0          # spread over the 400 nm box
    checked = 0
    for c in chains:
        assert list(c["type"]) == CHAIN
        _check_links(c)
        pos = c["pos"].astype(np.float64)
        if (pos <= 0).any() or (pos >= 400.0).any():
            continue                                        # clamped at the box
        for k in range(1, 8):
            step = pos[k - 1] - pos[k]
            want = radius[CHAIN[k] - 1] + radius[CHAIN[k]]
            assert abs(np.linalg.norm(step) - want) < 1e-4
            up = rotate(np.array([0.0, 1.0, 0.0]), c["rot"][k])
            assert np.allclose(step / want, up, atol=1e-4)
        checked += 1
    assert checked > 900


def test_line_and_complexes(scene):
    s, radius, complex1 = scene
    sizes, starts = layout(complex1)
    line = s.plain[starts[5]:starts[6]]
    base = np.array([200.0 - 1.8570869001109949, 200.0 + 2.8257686272096096, 200.0 - 18.748319527406373])
    want = base + np.outer(np.arange(8), [0.0, -1.5, 0.0])
    assert np.allclose(line["pos"], want, atol=1e-4)
    assert (line["rot"] == [0, 0, 0, 1]).all() and list(line["type"]) == CHAIN
    _check_links(line)
    for k, centre in ((4, [200.0, 200.0, 200.0]), (6, [200.0, 200.0, 120.0])):
        c = s.plain[starts[k]:starts[k + 1]]
        assert list(c["type"]) == [q["type"] for q in complex1]
        assert np.allclose(c["pos"], np.array(centre) + np.array([q["position"] for q in complex1]), atol=1e-4)
        assert np.allclose(c["rot"], [q["rotation"] for q in complex1])
        assert (c["n_bonds"] == 0).all()


def test_metabolite_seeding(scene):
    s, radius, complex1 = scene
    m = s.metabolic
    seeded = m["metabolites"][:, 0] == 100.0
    below = m["pos"][:, 0].astype(np.float64) < 400.0 * 0.2
    assert np.array_equal(seeded[1:], below[1:]) and not seeded[0]
