"""The benchmark workloads on the GPU, checked against the oracle where it
finishes in seconds and through size-independent properties at full size:

* C3 (8M particles, lipid shell + dense protein ball): the same structure
  scaled to a 160 nm cell, bit for bit against the oracle for 3 steps (long
  lists that spill into the pools); at full size ids kept, positions finite
  and in the box, no particle walking the stencil once the pools have grown,
  graph replay equal to kernel-by-kernel launches.
* C4 (256 cells, 2^24 particles in one batch): every instance keeps its
  count and ids; the two smallest instances and the median one equal the oracle run
  on their own documents.
* C5 (serving): requests drawn from C5's size and step distributions
  (scaled down), arriving while others run, with less capacity than the
  sum; the small ones equal the oracle, the others their cell run alone.
"""
import time

import numpy as np
import pytest

import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu


def _same(label, got, want):
    cg, pg, mg = got
    cw, pw, mw = want
    scenes.assert_same_state(label, cg, pg, mg, cw, pw, mw)


def _properties(label, before, after, L):
    c0, p0, m0 = before
    c1, p1, m1 = after
    for arr0, arr1, n0, n1 in ((p0, p1, c0.n_active, c1.n_active),
                               (m0, m1, c0.n_active_metabolic, c1.n_active_metabolic)):
        assert n1 == n0, label
        assert np.array_equal(np.sort(arr0["id"][:n0]), np.sort(arr1["id"][:n1])), label
        pos = arr1["pos"][:n1]
        assert np.isfinite(pos).all() and (pos >= 0.0).all() and (pos < L).all(), label
        assert np.isfinite(arr1["vel"][:n1]).all() and np.isfinite(arr1["rot"][:n1]).all(), label


# ---- C3 --------------------------------------------------------------------------

def test_c3_structure_scaled_matches_oracle():
    cfg, setup = scenes.c3_like_scene(0.08)
    print(f"[c3 scaled] {cfg['numParticles']} + {cfg['numMetabolicParticles']} particles, L {cfg['simSize']}",
          flush=True)
    sim = sc.Simulation(cfg, setup)
    sim.step(3)
    got = sim.snapshot()
    cnt = sim.export(sc.EXPORT_LIST_COUNTS, 0).reshape(-1, 2)
    sim.close()
    want = scenes.oracle_run(cfg, setup, 3)
    _same("c3 scaled, 3 steps", got, want)
    live = cnt[cnt[:, 0] != sc.LIST_WALKS, 0]
    assert live.max() > 16                    # the ball's lists spill past their slots into the pools


def test_c3_full_size_properties():
    cfg, setup = scenes.c3_scene()
    L = cfg["simSize"]
    print("[c3 full] graph", flush=True)
    a = sc.Simulation(cfg, setup)
    before = a.snapshot()
    for _ in range(3):             # pools grow before a step once the last step's demand has landed
        a.step(1)
        a.synchronize()
    after = a.snapshot()
    _properties("c3 full size", before, after, L)
    for q in (0, 1):
        cnt = a.export(sc.EXPORT_LIST_COUNTS, q).reshape(-1, 2)
        assert not (cnt == sc.LIST_WALKS).any(), "a particle still walks the stencil after the pools grew"
    a.close()
    del before
    print("[c3 full] eager", flush=True)
    b = sc.Simulation(cfg, setup)
    b.set_option("graph", 0)
    b.step(3)
    other = b.snapshot()
    b.close()
    _same("c3 graph vs eager", after, other)


# ---- C4 --------------------------------------------------------------------------

def test_c4_batch_instances_against_oracle():
    sizes = scenes.c4_sizes()
    docs = [scenes.c1_mix_scene(n, seed=42 + i) for i, n in enumerate(sizes)]
    t = time.perf_counter()
    batch = sc.Batch([c for c, _ in docs], [s for _, s in docs])
    print(f"[c4] batch of {len(docs)} ({sum(sizes)} particles) in {time.perf_counter() - t:.1f}s", flush=True)
    before = [batch.instance_counts(i) for i in range(len(docs))]
    firsts = {i: batch.instance_snapshot(i) for i in range(0, len(docs), 32)}
    batch.step(2)
    batch.synchronize()
    for i in range(len(docs)):
        c = batch.instance_counts(i)
        assert (c.n_active, c.n_active_metabolic, c.step) == (before[i].n_active, before[i].n_active_metabolic, 2)
    for i, snap in firsts.items():
        _properties(f"c4 instance {i}", snap, batch.instance_snapshot(i), docs[i][0]["simSize"])
    order = np.argsort(sizes)
    for i in (int(order[0]), int(order[1]), int(order[128])):
        print(f"[c4] instance {i}: {sizes[i]} particles vs oracle", flush=True)
        _same(f"c4 instance {i}", batch.instance_snapshot(i), scenes.oracle_run(*docs[i], 2))
    batch.close()


# ---- C5 --------------------------------------------------------------------------

def test_c5_request_mix_against_oracle():
    rng = np.random.default_rng(7)
    n = 10
    # C5: sizes log-uniform on [8192, 2097152], steps uniform on [100, 5000];
    # here both scaled down (sizes / 16, steps / 400) so that the run and the
    # oracle's checks take seconds
    sizes = np.exp(rng.uniform(np.log(512), np.log(131072), n)).astype(int)
    steps = rng.integers(1, 13, n)
    gaps = rng.exponential(1.0 / 20.0, n)
    docs = [scenes.c1_mix_scene(int(sizes[i]), seed=500 + i) for i in range(n)]
    cap_p = int(0.4 * sum(c["numParticles"] for c, _ in docs))
    cap_m = int(0.4 * sum(c["numMetabolicParticles"] for c, _ in docs))
    cap_p = max(cap_p, max(c["numParticles"] for c, _ in docs))
    cap_m = max(cap_m, max(c["numMetabolicParticles"] for c, _ in docs))
    doc = {"max_particles": cap_p, "max_metabolic_particles": cap_m, "max_instances": 8, "max_cells": 1 << 26}
    with sc.Server(*scenes.c1_mix_scene(8192, seed=0), doc) as srv:
        ids = []
        for i in range(n):                       # Poisson arrivals at 20 per second
            time.sleep(float(gaps[i]))
            ids.append(srv.submit(*docs[i], int(steps[i])))
        infos = [srv.wait(r) for r in ids]
        assert all(x.state == sc.REQ_DONE for x in infos)
        for i in range(n):
            got = srv.result(ids[i])
            if sizes[i] <= 17000:
                print(f"[c5] request {i}: {sizes[i]} particles, {steps[i]} steps vs oracle", flush=True)
                want = scenes.oracle_run(*docs[i], int(steps[i]))
            else:
                print(f"[c5] request {i}: {sizes[i]} particles, {steps[i]} steps vs alone", flush=True)
                sim = sc.Simulation(*docs[i])
                sim.step(int(steps[i]))
                want = sim.snapshot()
                sim.close()
            _same(f"c5 request {i}", got, want)
            srv.release(ids[i])
        assert sum(1 for x in sizes if x <= 17000) >= 3
