"""ORACLE -- test infrastructure only (parity checker and CPU baseline).

ctypes binding of liboracle.so, the serial host-C++ restatement of the
reference step.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LIB_OMP_PATH = os.path.join(HERE, "liboracle_omp.so")   # pure Jacobi passes over OpenMP threads
LIB_F64_PATH = os.path.join(HERE, "liboracle_f64.so")   # the same steps in double precision
LIB_PLANT_PATH = os.path.join(HERE, "liboracle_plant.so")   # with orc_plant(k): planted mistakes for tests
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "spatial-cell_amd"))

import spatial_cell as sc  # noqa: E402  (struct layouts of include/spatial_cell.h)

_P = ctypes.c_void_p
_libs = {}


def load(threaded=False, fp64=False, plants=False, path=None):
    path = path or (LIB_PLANT_PATH if plants else LIB_F64_PATH if fp64 else LIB_OMP_PATH if threaded else LIB_PATH)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with oracle/build.py")
    lib = ctypes.CDLL(path)
    lib.orc_create.restype = _P
    lib.orc_create.argtypes = [ctypes.POINTER(sc.Params), ctypes.POINTER(sc.Tables), ctypes.POINTER(sc.Counts),
                               _P, ctypes.c_size_t, _P, ctypes.c_size_t]
    lib.orc_destroy.argtypes = [_P]
    lib.orc_step.argtypes = [_P, ctypes.c_int]
    lib.orc_run_stage.restype = ctypes.c_int
    lib.orc_run_stage.argtypes = [_P, ctypes.c_int, ctypes.c_int]
    lib.orc_snapshot.argtypes = [_P, ctypes.POINTER(sc.Counts), _P, ctypes.c_size_t, _P, ctypes.c_size_t]
    lib.orc_debug_export.restype = ctypes.c_size_t
    lib.orc_debug_export.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_size_t]
    lib.orc_frame_size.restype = ctypes.c_size_t
    lib.orc_frame_size.argtypes = [_P]
    lib.orc_emit_frame.restype = ctypes.c_size_t
    lib.orc_emit_frame.argtypes = [_P, _P, ctypes.c_size_t]
    lib.orc_eval_primitive.argtypes = [ctypes.c_int, _P, _P, ctypes.c_size_t]
    lib.orc_philox.argtypes = [_P, _P, _P, ctypes.c_size_t]
    lib.orc_eval_math.argtypes = [ctypes.c_int, _P, _P, ctypes.c_size_t]
    if plants:
        lib.orc_plant.argtypes = [ctypes.c_int]
    _libs[path] = lib
    return lib


def eval_primitive(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    load().orc_eval_primitive(fn, x.ctypes.data, y.ctypes.data, len(x))
    return y


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    load().orc_philox(ctr.ctypes.data, key.ctypes.data, out.ctypes.data, len(ctr))
    return out


def eval_math(fn, a, b, t):
    a = np.asarray(a, dtype=np.float32).reshape(-1, 4)
    b = np.asarray(b, dtype=np.float32).reshape(-1, 4)
    t = np.asarray(t, dtype=np.float32).reshape(-1, 1)
    inp = np.ascontiguousarray(np.concatenate([a, b, t], axis=1), dtype=np.float32)
    out = np.zeros((len(inp), 4), dtype=np.float32)
    load().orc_eval_math(fn, inp.ctypes.data, out.ctypes.data, len(inp))
    return out


class Oracle:
    """threaded: the OpenMP build (same results, pure passes over threads).
    fp64: the double-precision build (every real quantity of the step a
    double; state read and reported as float).  plant: a planted mistake
    (orc_plant id of oracle.cpp) switched on in the planting build, for tests
    that show a check fails on it; 0 = none.  lib_path: another build of the
    same source (the CPU baseline's host-tuned one)."""
    def __init__(self, params, tables, counts, plain, met, threaded=False, fp64=False, plant=None, lib_path=None):
        self.lib = load(threaded, fp64, plants=plant is not None, path=lib_path)
        if plant is not None:
            self.lib.orc_plant(int(plant))
        self.params = params
        self._tables = sc.TablesHolder(tables)
        self._plain = np.ascontiguousarray(plain, dtype=sc.PARTICLE_DTYPE)
        self._met = np.ascontiguousarray(met, dtype=sc.PARTICLE_DTYPE)
        c = sc.Counts(*counts.as_tuple()) if hasattr(counts, "as_tuple") else sc.Counts(*counts)
        self.h = self.lib.orc_create(ctypes.byref(params), ctypes.byref(self._tables.struct), ctypes.byref(c),
                                     self._plain.ctypes.data, len(self._plain), self._met.ctypes.data, len(self._met))

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        self.lib.orc_step(self.h, n)

    def run_stage(self, stage, arg=0):
        if self.lib.orc_run_stage(self.h, stage, arg) != 0:
            raise ValueError(f"unknown stage {stage}")

    def snapshot(self):
        c = sc.Counts()
        plain = np.zeros(self.params.num_particles, dtype=sc.PARTICLE_DTYPE)
        met = np.zeros(self.params.num_metabolic_particles, dtype=sc.PARTICLE_DTYPE)
        self.lib.orc_snapshot(self.h, ctypes.byref(c), plain.ctypes.data, len(plain), met.ctypes.data, len(met))
        return c, plain, met

    def export(self, what, population):
        n = self.lib.orc_debug_export(self.h, what, population, None, 0)
        out = np.zeros(n // 4, dtype=np.uint32 if what in (sc.EXPORT_KEYS, sc.EXPORT_RANGES) else np.int32)
        self.lib.orc_debug_export(self.h, what, population, out.ctypes.data, n)
        return out

    def emit_frame(self):
        n = self.lib.orc_frame_size(self.h)
        b = ctypes.create_string_buffer(n)
        self.lib.orc_emit_frame(self.h, b, n)
        return b.raw
