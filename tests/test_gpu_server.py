"""Serving loop (sc_server_*): a request's result equals running its
documents alone for its steps, whether it was admitted next to others, while
others were running, or after waiting for capacity."""
import pytest

import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu


def _alone(cfg, setup, steps):
    sim = sc.Simulation(cfg, setup)
    sim.step(steps)
    out = sim.snapshot()
    sim.close()
    return out


def _same(label, got, want):
    cg, pg, mg = got
    cw, pw, mw = want
    scenes.assert_same_state(label, cg, pg, mg, cw, pw, mw)


def _template():
    return scenes.mixed_scene(scale=0.3, L=30.0, seed=0, init=True)


def test_server_results_equal_alone():
    reqs = [(scenes.mixed_scene(scale=0.5, L=40.0, seed=11, init=True), 4),
            (scenes.mixed_scene(scale=1.0, L=60.0, seed=12, init=True), 7),
            (scenes.mixed_scene(scale=0.3, L=30.0, seed=13, init=True), 2),
            (scenes.mixed_scene(scale=0.6, L=50.0, seed=14, init=True), 0)]
    cap_p = sum(c["numParticles"] for (c, _), _ in reqs)
    cap_m = sum(c["numMetabolicParticles"] for (c, _), _ in reqs)
    doc = {"max_particles": cap_p, "max_metabolic_particles": cap_m, "max_instances": 8, "max_cells": 1 << 16}
    with sc.Server(*_template(), doc) as srv:
        ids = [srv.submit(c, s, n) for (c, s), n in reqs[:2]]
        srv.wait(ids[0])                                   # the second one is still running
        ids += [srv.submit(c, s, n) for (c, s), n in reqs[2:]]
        for k, (rid, ((c, s), n)) in enumerate(zip(ids, reqs)):
            info = srv.wait(rid)
            assert info.state == sc.REQ_DONE and info.steps_done == n
            assert info.t_submit <= info.t_admit <= info.t_done
            print(f"[server] request {k}: {n} steps", flush=True)
            got = srv.result(rid)
            _same(f"request {k}", got, _alone(c, s, n))
            if k in (0, 2):                                 # and against the oracle itself
                _same(f"request {k} vs oracle", got, scenes.oracle_run(c, s, n))
            srv.release(rid)
        st = srv.stats()
        assert st["admitted"] == 4 and st["retired"] == 4
        assert st["steps"] >= 7 and st["slot_steps"] > 0 and st["t_step"] > 0.0


def test_server_queues_for_capacity_and_refuses_oversize():
    reqs = [(scenes.mixed_scene(scale=1.0, L=60.0, seed=21 + k, init=True), 3 + k) for k in range(3)]
    big = max(c["numParticles"] for (c, _), _ in reqs)
    bigm = max(c["numMetabolicParticles"] for (c, _), _ in reqs)
    doc = {"max_particles": big, "max_metabolic_particles": bigm, "max_instances": 4, "max_cells": 1 << 16}
    with sc.Server(*_template(), doc) as srv:
        ids = [srv.submit(c, s, n) for (c, s), n in reqs]    # only one fits at a time
        infos = [srv.wait(rid) for rid in ids]
        for k in range(1, 3):                                 # first come, first served, one after another
            assert infos[k].t_admit >= infos[k - 1].t_done
        for k, (rid, ((c, s), n)) in enumerate(zip(ids, reqs)):
            _same(f"queued request {k}", srv.result(rid), _alone(c, s, n))
        c, s = scenes.mixed_scene(scale=2.0, L=80.0, seed=5, init=True)
        with pytest.raises(sc.SpatialCellError) as e:
            srv.submit(c, s, 1)
        assert e.value.status == 4                          # SC_ERR_BUDGET
        c, s = scenes.mixed_scene(scale=0.3, L=30.0, seed=6, init=True)
        with pytest.raises(sc.SpatialCellError):
            srv.submit(dict(c, relaxationSteps=8), s, 1)       # constants differ from the template


def test_server_failed_retire_frees_its_slot():
    """A finished cell whose final state cannot be staged (the retire stage
    limit, forced here) fails, but still leaves the batch: with one instance
    slot, the next request is admitted into that slot and finishes, equal to
    the oracle."""
    big = (scenes.mixed_scene(scale=3.0, L=90.0, seed=31, init=True), 2)
    small = (scenes.mixed_scene(scale=0.3, L=30.0, seed=32, init=True), 3)
    doc = {"max_particles": big[0][0]["numParticles"], "max_metabolic_particles": big[0][0]["numMetabolicParticles"],
           "max_instances": 1, "max_cells": 1 << 16, "max_retire_stage_bytes": 1 << 20}
    with sc.Server(*_template(), doc) as srv:
        r0 = srv.submit(*big[0], big[1])
        r1 = srv.submit(*small[0], small[1])
        with pytest.raises(sc.SpatialCellError):
            srv.wait(r0)
        assert srv.poll(r0).state == sc.REQ_FAILED
        info = srv.wait(r1)
        assert info.state == sc.REQ_DONE and info.steps_done == small[1]
        (c, s), n = small
        _same("request after a failed retire", srv.result(r1), scenes.oracle_run(c, s, n))
        st = srv.stats()
        assert st["admitted"] == 2 and st["retired"] == 2
