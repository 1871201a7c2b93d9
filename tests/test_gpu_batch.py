"""Batches (sc_batch_create): every instance of a batch evolves exactly as
the same two documents run as a single cell -- same slots in the same
order, same counts, same frames -- whatever the other instances are; some
instances are also compared with the oracle run on their own documents."""
import numpy as np
import pytest

import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu


def _instances(boundary):
    return [scenes.mixed_scene(scale=0.5, L=40.0, boundary=boundary, seed=1, init=True),
            scenes.mixed_scene(scale=1.0, L=60.0, boundary=boundary, seed=2, init=True),
            scenes.mixed_scene(scale=0.3, L=30.0, boundary=boundary, seed=3, init=True)]


def _same(label, a, b):
    ca, pa, ma = a
    cb, pb, mb = b
    assert (ca.n_active, ca.n_active_metabolic, ca.next_particle_id, ca.step) == \
           (cb.n_active, cb.n_active_metabolic, cb.next_particle_id, cb.step), label
    scenes.assert_same_state(label, ca, pa, ma, cb, pb, mb)


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_batch_instances_equal_single_cells(boundary):
    inst = _instances(boundary)
    print(f"[batch {boundary}] {len(inst)} instances", flush=True)
    batch = sc.Batch([c for c, _ in inst], [s for _, s in inst])
    alone = [sc.Simulation(c, s) for c, s in inst]
    for i, sim in enumerate(alone):
        _same(f"{boundary}/init/{i}", batch.instance_snapshot(i), sim.snapshot())
    for step in range(3):
        batch.step(1)
        for i, sim in enumerate(alone):
            sim.step(1)
            _same(f"{boundary}/step{step}/{i}", batch.instance_snapshot(i), sim.snapshot())
    for i, sim in enumerate(alone):
        assert batch.instance_frame_header(i) == sim.frame_header()
        assert batch.instance_emit_frame(i) == sim.emit_frame()
        c = batch.instance_counts(i)
        assert c.step == 3
    for i in (0, 2):                     # and against the oracle on the instance's own documents
        _same(f"{boundary}/oracle/{i}", batch.instance_snapshot(i), scenes.oracle_run(*inst[i], 3))


def test_batch_tiny_periodic_instances_of_different_sizes():
    """Instances of a few dozen particles each, so that one wave of 64 slots
    holds the low corner of several boxes of different sizes: lanes of two
    instances then share a column index whose periodic image shift differs,
    and each must still see its own box's images."""
    inst = [scenes.mixed_scene(scale=s, L=L, boundary="periodic", seed=10 + k, init=True)
            for k, (s, L) in enumerate([(0.05, 30.0), (0.06, 50.0), (0.05, 40.0), (0.07, 60.0), (0.05, 30.0)])]
    print(f"[batch tiny periodic] {len(inst)} instances", flush=True)
    batch = sc.Batch([c for c, _ in inst], [s for _, s in inst])
    alone = [sc.Simulation(c, s) for c, s in inst]
    for step in range(3):
        batch.step(1)
        for i, sim in enumerate(alone):
            sim.step(1)
            _same(f"tiny/step{step}/{i}", batch.instance_snapshot(i), sim.snapshot())


def test_batch_reactions_happen_per_instance():
    """The reaction scene binds and polymerises inside a batch too (ids and
    capacities are per instance): counts grow exactly as alone."""
    inst = [scenes.mixed_scene(scale=1.0, L=60.0, seed=7, init=True),
            scenes.mixed_scene(scale=0.6, L=50.0, seed=8, init=True)]
    batch = sc.Batch([c for c, _ in inst], [s for _, s in inst])
    alone = [sc.Simulation(c, s) for c, s in inst]
    batch.step(12)
    grew = False
    for i, sim in enumerate(alone):
        sim.step(12)
        a, b = batch.instance_snapshot(i), sim.snapshot()
        _same(f"reactions/{i}", a, b)
        grew = grew or int((a[1]["n_bonds"] > 0).sum()) > 0
    assert grew
    _same("reactions/1 vs oracle", batch.instance_snapshot(1), scenes.oracle_run(*inst[1], 12))


def test_batch_rejects_mismatched_shared_constants():
    c0, s0 = scenes.mixed_scene(scale=0.3, L=30.0, seed=1)
    c1, s1 = scenes.mixed_scene(scale=0.3, L=30.0, seed=2)
    c1 = dict(c1, relaxationSteps=16)
    with pytest.raises(sc.SpatialCellError) as e:
        sc.Batch([c0, c1], [s0, s1])
    assert "sub-iteration" in str(e.value)
    b = sc.Batch([c0], [s0])
    with pytest.raises(sc.SpatialCellError):
        b.snapshot()                     # single-cell call on a batch handle
