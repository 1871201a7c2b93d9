"""The command-line driver end to end on the GPU: its frames.dat must hold
the header and, bit for bit, the frames the library emits for the same run
(frame 0 after the overlap removal, then one after every
persistEveryNthFrame-th step: main.cu:315-323, 666-680), written through the
asynchronous double-buffered path."""
import json
import os
import subprocess

import numpy as np
import pytest

import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spatial-cell_amd", "spatial_cell_sim")


def test_driver_frames_match_the_library(tmp_path):
    cfg, setup = scenes.c1_scene(init=True)
    cfg = dict(cfg, steps=5, persistEveryNthFrame=2, reportEveryNthFrame=2)
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    (tmp_path / "setup.json").write_text(json.dumps(setup))
    (tmp_path / "results").mkdir()
    print("[driver] c1 scene, 5 steps, a frame every 2", flush=True)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "[Simulating...]" in r.stdout and "step 4, nActiveParticles" in r.stdout and "Done" in r.stdout
    data = (tmp_path / "results" / "frames.dat").read_bytes()

    # the same run through the library, synchronous frames
    sim = sc.Simulation(cfg, dict(setup, init=False))
    want = [sim.frame_header()]
    sim.run_stage(sc.STAGE_SORT, 0)
    sim.run_stage(sc.STAGE_REMOVE_INTERFERING, 0)
    want.append(sim.emit_frame())
    sim.run_stage(sc.STAGE_COMPLEXIFY, -1)
    for step in range(5):
        sim.step(1)
        if step % 2 == 0:
            want.append(sim.emit_frame())
    sim.close()
    frame = len(want[1])
    want = b"".join(want)
    assert len(want) == 12 + 4 * frame               # header, frame 0, after steps 0, 2 and 4
    assert len(data) == len(want) and data == want


def test_async_frames_equal_synchronous_frames():
    cfg, setup = scenes.c1_scene(init=True)
    a = sc.Simulation(cfg, setup)
    got = a.emit_frames_async(2, 3)
    a.close()
    b = sc.Simulation(cfg, setup)
    want = [b.emit_frame()]
    for _ in range(2):
        b.step(2)
        want.append(b.emit_frame())
    b.close()
    assert len(got) == 3 and all(x == y for x, y in zip(got, want))


def test_init_through_stages_equals_create_with_init():
    """The driver's init (sort, overlap removal, frame 0, complexify at step
    -1 through sc_run_stage) leaves the state sc_create's init does."""
    cfg, setup = scenes.c1_scene(init=True)
    a = sc.Simulation(cfg, setup)
    b = sc.Simulation(cfg, dict(setup, init=False))
    for stage, arg in ((sc.STAGE_SORT, 0), (sc.STAGE_REMOVE_INTERFERING, 0), (sc.STAGE_COMPLEXIFY, -1)):
        b.run_stage(stage, arg)
    for s in (a, b):
        s.step(2)
    ca, pa, ma = a.snapshot()
    cb, pb, mb = b.snapshot()
    scenes.assert_same_state("init", ca, pa, ma, cb, pb, mb)
