"""sc_bin_noalloc (binning with its workspace in the library's fixed device
array, no allocation) against sc_bin on two calls in a row, the second with
a smaller input on another grid, bit for bit; and its size limits."""
import numpy as np
import pytest

import hipmem as hm
import scenes
import spatial_cell as sc
from test_gpu_bin import _arrays

pytestmark = pytest.mark.gpu


def _bin(pos, hdr, L, cell, boundary, noalloc):
    n = len(pos)
    cells = sc.grid_cells(L, cell)
    tp = hm.DeviceArray(n * 4, np.float32, init=pos.reshape(-1))
    th = hm.DeviceArray(n * 4, np.int32, init=hdr.reshape(-1))
    keys = hm.DeviceArray(n, np.uint32, fill_byte=0x5A)
    perm = hm.DeviceArray(n, np.int32, fill_byte=0xA5)
    cs = hm.DeviceArray(cells + 1, np.uint32, fill_byte=0x7B)
    sc.bin_arrays(tp.ptr, th.ptr, n, L, cell, boundary, keys.ptr, perm.ptr, cs.ptr, noalloc=noalloc)
    hm.synchronize()
    return keys.get(), perm.get(), cs.get()


def _case(n, L, boundary, seed, inactive):
    cfg, setup = scenes.uniform_scene(n, L, boundary=boundary, seed=seed, init=False)
    plain = sc.Setup(cfg, setup).plain.copy()
    rng = np.random.default_rng(seed)
    plain["flags"][rng.random(len(plain)) < inactive] = 0
    pos, hdr = _arrays(plain)
    return pos, hdr, cfg["simSize"], cfg["gridCellSize"], \
        sc.BOUNDARY_PERIODIC if boundary == "periodic" else sc.BOUNDARY_CLAMP


def test_noalloc_equals_sc_bin_twice_second_smaller():
    big = _case(1 << 19, 700.0, "periodic", 5, 0.1)     # 235,929 plain slots, a 128^3 grid
    small = _case(20000, 300.0, "clamp", 6, 0.3)
    for label, (pos, hdr, L, cell, b) in (("first, larger", big), ("second, smaller", small)):
        print(f"[noalloc] {label}: {len(pos)} slots, L {L}", flush=True)
        want = _bin(pos, hdr, L, cell, b, noalloc=False)
        got = _bin(pos, hdr, L, cell, b, noalloc=True)
        for name, x, y in zip(("keys", "perm", "cell_start"), got, want):
            assert np.array_equal(x, y), (label, name)


def test_noalloc_limits():
    x = hm.DeviceArray(16, np.float32)
    cs = hm.DeviceArray(16, np.uint32)
    with pytest.raises(sc.SpatialCellError) as e:
        sc.bin_arrays(x.ptr, x.ptr, sc.BIN_NOALLOC_MAX_SLOTS + 1, 400.0, 10.0, 0, x.ptr, x.ptr, cs.ptr,
                      noalloc=True)
    assert e.value.status == 4 and "SC_BIN_NOALLOC_MAX_SLOTS" in str(e.value)
    with pytest.raises(sc.SpatialCellError) as e:
        sc.bin_arrays(x.ptr, x.ptr, 4, 5000.0, 10.0, 0, x.ptr, x.ptr, cs.ptr, noalloc=True)
    assert e.value.status == 4 and "SC_BIN_NOALLOC_MAX_CELLS" in str(e.value)
