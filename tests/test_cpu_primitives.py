"""Pins the scalar primitives shared by the oracle and the kernels: accuracy
against numpy float64, exact special values, Philox4x32-10 against rocRAND's
engine and the published known-answer vectors, float->binary16 against
numpy's round-to-nearest-even, and the agreement of the g++ (oracle) and clang
(library host pass) builds of the same code."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as orc
import spatial_cell as sc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ulp_err(y, ref):
    y = y.astype(np.float64)
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(y - ref) / sp


@pytest.mark.parametrize("fn,lo,hi,ref,maxulp", [
    (0, -6.3, 6.3, np.sin, 2.0),
    (1, -6.3, 6.3, np.cos, 2.0),
    (2, -1.0, 1.0, np.arccos, 2.0),
    (3, 1e-6, 100.0, np.log, 2.0),
    (4, 1e-6, 100.0, np.log2, 3.0),
    (5, -20.0, 20.0, np.exp2, 2.0),
    (6, 1e-3, 1.0, lambda x: np.power(x, np.float64(np.float32(0.2))), 3.0),
])
def test_transcendental_accuracy(fn, lo, hi, ref, maxulp):
    x = np.random.default_rng(fn).uniform(lo, hi, 200000).astype(np.float32)
    y = orc.eval_primitive(fn, x)
    r = ref(x.astype(np.float64))
    # near zeros of sin/cos the relative ulp is meaningless: use an absolute floor
    err = np.abs(y.astype(np.float64) - r) / np.maximum(np.spacing(np.abs(r).astype(np.float32)), 2e-8)
    assert err.max() <= maxulp, (fn, float(err.max()), x[np.argmax(err)])


def test_special_values():
    f = orc.eval_primitive
    assert f(0, [0.0])[0] == 0.0 and f(1, [0.0])[0] == 1.0
    assert f(2, [1.0])[0] == 0.0
    assert f(2, [-1.0])[0] == np.float32(np.pi)
    assert f(2, [0.0])[0] == np.float32(np.pi / 2)
    assert np.isnan(f(2, [1.0000001, -1.0000001, np.nan])).all()
    assert f(3, [1.0])[0] == 0.0 and f(4, [1.0])[0] == 0.0
    assert f(4, [8.0])[0] == 3.0 and f(4, [0.25])[0] == -2.0
    k = np.arange(-126, 128, dtype=np.float32)
    assert np.array_equal(f(5, k), np.exp2(k.astype(np.float64)).astype(np.float32))
    assert f(6, [1.0])[0] == 1.0 and f(6, [0.0])[0] == 0.0
    assert np.isnan(f(6, [-1.0])[0])
    x = np.random.default_rng(3).uniform(0, 1e4, 10000).astype(np.float32)
    assert np.array_equal(f(7, x), np.sqrt(x))     # IEEE correctly rounded


def test_sin_cos_identity_and_quadrants():
    x = np.linspace(-6.28, 6.28, 100001).astype(np.float32)
    s, c = orc.eval_primitive(0, x), orc.eval_primitive(1, x)
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 3e-7
    assert np.array_equal(np.sign(s[np.abs(s) > 1e-3]), np.sign(np.sin(x[np.abs(s) > 1e-3].astype(np.float64))))


def test_f16_round_to_nearest_even_matches_numpy():
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-70000, 70000, 300000), rng.uniform(-1e-3, 1e-3, 100000),
        rng.standard_normal(100000) * 1e-6,
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 1e9, np.inf, -np.inf, 5.960464477539063e-08,
         2.9802322387695312e-08, 2.98023259e-08, 6.103515625e-05, 6.1035153e-05],
    ]).astype(np.float32)
    # exact ties between halves
    halves = rng.integers(0, 0x7bff, 100000).astype(np.uint16).view(np.float16).astype(np.float32)
    nxt = (rng.integers(0, 0x7bff, 100000).astype(np.uint16) + 1).view(np.float16).astype(np.float32)
    x = np.concatenate([x, ((halves.astype(np.float64) + nxt.astype(np.float64)) / 2).astype(np.float32)])
    bits = orc.eval_primitive(8, x).view(np.uint32).astype(np.uint16)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(bits, ref)
    nan_bits = orc.eval_primitive(8, np.array([np.nan], np.float32)).view(np.uint32)[0]
    assert (nan_bits & 0x7c00) == 0x7c00 and (nan_bits & 0x3ff) != 0


RANDOM123_KAT = [  # Philox4x32-10 known answers published with Random123 (Salmon et al., SC'11)
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff, 0xffffffff], [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_known_answers():
    for ctr, key, out in RANDOM123_KAT:
        assert orc.philox([ctr], [key])[0].tolist() == out


def test_philox_matches_rocrand_engine():
    with open(os.path.join(GOLD, "philox_rocrand.json")) as f:
        rows = json.load(f)["blocks"]
    ctr = np.array([r["counter"] for r in rows], dtype=np.uint32)
    key = np.array([r["key"] for r in rows], dtype=np.uint32)
    out = np.array([r["out"] for r in rows], dtype=np.uint32)
    assert np.array_equal(orc.philox(ctr, key), out)


def test_uniform_and_normal_moments():
    x = np.random.default_rng(2).integers(0, 2**32, 400000, dtype=np.uint32).view(np.float32)
    n = orc.eval_primitive(9, x).astype(np.float64)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1.0) < 0.01
    assert abs((n ** 4).mean() - 3.0) < 0.05


def test_shoemake_rotation_is_uniform():
    # E[R(q) y] = 0 and E[(R(q) y . e)^2] = 1/3 for q uniform on S3
    rng = np.random.default_rng(5)
    u = rng.random((100000, 3)).astype(np.float32)
    a = np.concatenate([u, np.zeros((len(u), 1), np.float32)], axis=1)
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]), np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]), np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], 1)
    up = np.tile(np.array([0, 1, 0, 0], np.float32), (len(u), 1))
    v = orc.eval_math(0, up, q.astype(np.float32), np.zeros(len(u)))[:, :3].astype(np.float64)
    assert np.abs(v.mean(axis=0)).max() < 0.01
    assert np.abs((v ** 2).mean(axis=0) - 1 / 3).max() < 0.01
    assert np.abs(np.linalg.norm(v, axis=1) - 1).max() < 1e-6
    del a


def test_gcc_and_clang_builds_agree():
    lib = sc.load()
    rng = np.random.default_rng(17)
    for fn in range(12):
        x = rng.uniform(-50, 50, 50000).astype(np.float32)
        if fn in (3, 4, 6):
            x = np.abs(x) + np.float32(1e-6)
        if fn == 2:
            x = rng.uniform(-1, 1, 50000).astype(np.float32)
        y_clang = np.zeros_like(x)
        assert lib.sc_eval_primitive(0, fn, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     y_clang.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(x), 0) == 0
        y_gcc = orc.eval_primitive(fn, x)
        same = (y_clang.view(np.uint32) == y_gcc.view(np.uint32)) | (np.isnan(y_clang) & np.isnan(y_gcc))
        assert same.all(), (fn, int((~same).sum()))


def test_contact_free_updates_reach_a_fixed_point():
    """The relax, metabolic relax and diffusion passes skip particles without
    list entries once their values are a fixed point of the per-pass update.
    With no terms the update is v -> v + 0/1 (velocity, metabolites) and
    q -> qmul(slerp(id, id, 1), q) (angular velocity).  Every product in the
    latter is 1*x or 0*x, so the result depends only on the sign/zero class of
    each component: enumerating all 4^4 classes proves that two applications
    reach the fixed point (passes skipped from the third / fourth on)."""
    f32 = np.float32
    vals = [f32(0.0), f32(-0.0), f32(1.5), f32(-2.0)]

    def qmul(a, b):
        cx = a[1] * b[2] - a[2] * b[1]
        cy = a[2] * b[0] - a[0] * b[2]
        cz = a[0] * b[1] - a[1] * b[0]
        d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
        return (a[0] * b[3] + b[0] * a[3] + cx, a[1] * b[3] + b[1] * a[3] + cy,
                a[2] * b[3] + b[2] * a[3] + cz, a[3] * b[3] - d)

    # slerp(id, id, 1): c = 1 takes the linear branch, s1 = 0, s2 = 1
    s1, s2 = f32(1.0) - f32(1.0), f32(1.0)
    ident = tuple(s1 * x + s2 * x for x in (f32(0), f32(0), f32(0), f32(1)))
    assert [np.float32(x).tobytes() for x in ident] == [f32(0).tobytes()] * 3 + [f32(1).tobytes()]
    import itertools
    needed = 0
    for p in itertools.product(vals, repeat=4):
        seq = [tuple(f32(x) for x in p)]
        for _ in range(3):
            seq.append(tuple(f32(x) for x in qmul(ident, seq[-1])))
        b = [b"".join(np.float32(x).tobytes() for x in q) for q in seq]
        assert b[3] == b[2], p                      # g^3 == g^2 for every class
        needed = max(needed, 2 if b[2] != b[1] else (1 if b[1] != b[0] else 0))
    assert needed == 2                               # and two applications are needed in general
    zero = f32(0.0) / max(f32(0.0), f32(1.0))
    for v in vals:
        once = f32(v + zero)
        assert f32(once + zero).tobytes() == once.tobytes()
