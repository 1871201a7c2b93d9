"""sc_bin (updateGridAndSort over caller-owned device arrays,
spatial_cell_sim/src/grid.cuh:111-139) against the oracle's sort stage: the
sorted keys (inactive 0xFFFFFFFF), the stable permutation and the cell table
bit for bit, on scenes with inactive slots, particles on the box faces, all
particles in one cell, an empty array, and a 2^18-particle cell."""
import numpy as np
import pytest

import hipmem as hm
import oracle as orc
import scenes
import spatial_cell as sc

pytestmark = pytest.mark.gpu


def _arrays(plain):
    pos = np.zeros((len(plain), 4), np.float32)
    pos[:, :3] = plain["pos"]
    pos[:, 3] = plain["radius"]
    hdr = np.stack([plain["id"], plain["type"], plain["flags"], plain["hydrophobic"]], axis=1).astype(np.int32)
    return pos, hdr


def run_bin(pos, hdr, L, cell, boundary, fn=None):
    n = len(pos)
    cells = sc.grid_cells(L, cell)
    tp = hm.DeviceArray(n * 4, np.float32, init=pos.reshape(-1))
    th = hm.DeviceArray(n * 4, np.int32, init=hdr.reshape(-1))
    keys = hm.DeviceArray(n, np.uint32, fill_byte=0x5A)
    perm = hm.DeviceArray(n, np.int32, fill_byte=0xA5)
    cs = hm.DeviceArray(cells + 1, np.uint32, fill_byte=0x7B)
    (fn or sc.bin_arrays)(tp.ptr, th.ptr, n, L, cell, boundary, keys.ptr, perm.ptr, cs.ptr)
    hm.synchronize()
    return keys.get(), perm.get(), cs.get()


def oracle_bin(cfg, setup, plain):
    s = sc.Setup(cfg, setup)
    counts = sc.Counts(len(plain), 0, len(plain), 0)
    o = orc.Oracle(s.params, s.tables, counts, plain, np.zeros(0, sc.PARTICLE_DTYPE))
    o.run_stage(sc.STAGE_SORT)
    return o.export(sc.EXPORT_KEYS, 0), o.export(sc.EXPORT_PERMUTATION, 0), o.export(sc.EXPORT_RANGES, 0)


def check_against_oracle(cfg, setup, plain, fn=None):
    L, cell = cfg["simSize"], cfg["gridCellSize"]
    boundary = sc.BOUNDARY_PERIODIC if cfg.get("boundary") == "periodic" else sc.BOUNDARY_CLAMP
    pos, hdr = _arrays(plain)
    keys, perm, cs = run_bin(pos, hdr, L, cell, boundary, fn)
    ok, op, orng = oracle_bin(cfg, setup, plain)
    n = len(plain)
    assert np.array_equal(keys, ok[:n]), "keys"
    assert np.array_equal(perm, op[:n]), "permutation"
    # the CSR table as the reference's {start, end} of every non-empty cell
    cells = sc.grid_cells(L, cell)
    assert len(cs) == cells + 1
    assert np.all(np.diff(cs.astype(np.int64)) >= 0)
    n_act = int((hdr[:, 2] & 1).sum()) if n else 0
    assert cs[0] == 0 and cs[cells] == n_act
    ne = np.nonzero(np.diff(cs.astype(np.int64)))[0]
    got = np.stack([ne, cs[ne], cs[ne + 1]], axis=1).astype(np.uint32).reshape(-1)
    assert np.array_equal(got, orng), "cell ranges"
    return keys, perm, cs


@pytest.mark.parametrize("boundary", ["clamp", "periodic"])
def test_bin_matches_oracle_with_inactive_slots(boundary):
    cfg, setup = scenes.mixed_scene(seed=11, boundary=boundary)
    plain = sc.Setup(cfg, setup).plain.copy()
    rng = np.random.default_rng(2)
    plain["flags"][rng.random(len(plain)) < 0.15] = 0
    print(f"[bin] mixed {boundary}: {len(plain)} slots", flush=True)
    keys, _, _ = check_against_oracle(cfg, setup, plain)
    assert (keys == 0xFFFFFFFF).sum() == (plain["flags"] & 1 == 0).sum()


def test_bin_faces_one_cell_and_c1_size():
    cfg, setup = scenes.c1_scene(boundary="clamp")
    plain = sc.Setup(cfg, setup).plain.copy()
    L = cfg["simSize"]
    plain["pos"][:64] = [L, L, L]             # the last pow2 cell index below n - 1 (config.cuh:66-74)
    plain["pos"][64:128] = [0.0, 0.0, 0.0]
    plain["pos"][128:140, 0] = L
    print(f"[bin] c1 clamp with faces: {len(plain)} slots", flush=True)
    check_against_oracle(cfg, setup, plain)
    one = plain[:3000].copy()
    one["pos"] = [123.4, 55.5, 301.0]          # every slot in one cell: the permutation is the identity
    _, perm, _ = check_against_oracle(cfg, setup, one)
    assert np.array_equal(perm, np.arange(len(one)))


def test_bin_empty_and_all_inactive():
    cfg, setup = scenes.mixed_scene(seed=3)
    cells = sc.grid_cells(cfg["simSize"], cfg["gridCellSize"])
    z4 = np.zeros((0, 4), np.float32)
    keys, perm, cs = run_bin(z4, z4.astype(np.int32), cfg["simSize"], cfg["gridCellSize"], sc.BOUNDARY_CLAMP)
    assert len(keys) == 0 and np.all(cs == 0) and len(cs) == cells + 1
    plain = sc.Setup(cfg, setup).plain.copy()
    plain["flags"] = 0
    keys, perm, cs = check_against_oracle(cfg, setup, plain)
    assert np.all(keys == 0xFFFFFFFF) and np.array_equal(perm, np.arange(len(plain))) and np.all(cs == 0)


def test_bin_large_uniform_cell():
    cfg, setup = scenes.uniform_scene(1 << 18, 630.0, boundary="periodic", init=False)
    plain = sc.Setup(cfg, setup).plain.copy()
    print(f"[bin] uniform periodic: {len(plain)} slots", flush=True)
    check_against_oracle(cfg, setup, plain)


def test_bin_rejects_bad_arguments():
    x = hm.DeviceArray(16, np.float32)
    cs = hm.DeviceArray(sc.grid_cells(60.0, 10.0) + 1, np.uint32)
    with pytest.raises(sc.SpatialCellError) as e:
        sc.bin_arrays(x.ptr + 4, x.ptr, 4, 60.0, 10.0, 0, x.ptr, x.ptr, cs.ptr)
    assert e.value.status == 1 and "aligned" in str(e.value)
    with pytest.raises(sc.SpatialCellError):
        sc.bin_arrays(x.ptr, x.ptr, -1, 60.0, 10.0, 0, x.ptr, x.ptr, cs.ptr)
    with pytest.raises(sc.SpatialCellError):
        sc.bin_arrays(x.ptr, x.ptr, 4, 60.0, 0.0, 0, x.ptr, x.ptr, cs.ptr)
