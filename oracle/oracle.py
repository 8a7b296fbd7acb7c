"""ctypes wrapper of the C oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module. It is the
checker for the MI355X path, never part of it.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

CELLTYPE, GRAINID, CRITTIMESTEP, LAYERID, UNDERCOOLING, UCHANGE, MELTED, DIAGLEN, DOCENTER, CRITDIAG, \
    NUCLEI_LOC, NUCLEI_TIME = range(12)
# planted mistakes in CellCapture's pass 2 (oracle_set_mutation)
MUT_NONE, MUT_DESCENDING, MUT_LCOUNT, MUT_GROW_NEW = range(4)

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(str(LIB))
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
        L.oracle_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.oracle_load.restype = ctypes.c_int
        L.oracle_load_seeded.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.oracle_load_seeded.restype = ctypes.c_int
        L.oracle_dims.argtypes = [vp, i64p]
        L.oracle_params.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_field.argtypes = [vp, ctypes.c_int, vp]
        L.oracle_get_field.restype = ctypes.c_int
        L.oracle_step.argtypes = [vp, ctypes.c_int]
        L.oracle_check.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.oracle_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_write_outputs.argtypes = [vp, ctypes.c_char_p]
        L.oracle_write_outputs.restype = ctypes.c_int
        L.oracle_free.argtypes = [vp]
        L.oracle_next_layer.argtypes = [vp]
        L.oracle_next_layer.restype = ctypes.c_int
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_mt64_draws.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_uniform_draws.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_minstd_draws.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_normal_draws.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)]
        L.oracle_seed_cdl.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.oracle_capture_octahedron.argtypes = [ctypes.POINTER(ctypes.c_float)] + [ctypes.c_double] * 6 + [
            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.oracle_growth_increment.argtypes = [ctypes.c_double] * 4 + [ctypes.c_float]
        L.oracle_growth_increment.restype = ctypes.c_double
        L.oracle_set_field.argtypes = [vp, ctypes.c_int, vp]
        L.oracle_set_field.restype = ctypes.c_int
        L.oracle_set_orientation.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.oracle_set_orientation.restype = ctypes.c_int
        L.oracle_set_mutation.argtypes = [vp, ctypes.c_int]
        L.oracle_set_threads.argtypes = [vp, ctypes.c_int]
        L.oracle_load_state.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int,
                                        vp, vp, ctypes.POINTER(vp)]
        L.oracle_load_state.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class OracleSim:
    """One deck, initialised and stepped serially on the CPU (reference Serial semantics)."""

    def __init__(self, deck, data_dir, seed: int = 0):
        self._h = ctypes.c_void_p()
        if lib().oracle_load_seeded(str(deck).encode(), str(data_dir).encode(), seed, ctypes.byref(self._h)) != 0:
            raise RuntimeError("oracle: " + lib().oracle_last_error().decode())
        self.deck = str(deck)

    @classmethod
    def from_state(cls, deck, data_dir, cycle: int, ct_raw, gid, dl, centre, nloc, ntime) -> "OracleSim":
        """A type-C simulation started from another run's state at the end of step `cycle`:
        ct_raw (uint8: type | 0x10 for cell-centred octahedra), GrainID (int32), DiagonalLength
        (float32) and centres (float64, 3 per cell) of every cell, and the nucleation sites."""
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        arrs = [np.ascontiguousarray(ct_raw, dtype=np.uint8), np.ascontiguousarray(gid, dtype=np.int32),
                np.ascontiguousarray(dl, dtype=np.float32), np.ascontiguousarray(centre, dtype=np.float64),
                np.ascontiguousarray(nloc, dtype=np.int32), np.ascontiguousarray(ntime, dtype=np.int32)]
        if not (arrs[0].size == arrs[1].size == arrs[2].size == arrs[3].size // 3 and arrs[4].size == arrs[5].size):
            raise ValueError("inconsistent state arrays")
        ct, g, d, c, nl, nt = arrs
        if lib().oracle_load_state(str(deck).encode(), str(data_dir).encode(), cycle, ct.ctypes.data, g.ctypes.data,
                                   d.ctypes.data, c.ctypes.data, nl.size, nl.ctypes.data, nt.ctypes.data,
                                   ctypes.byref(self._h)) != 0:
            raise RuntimeError("oracle: " + lib().oracle_last_error().decode())
        self.deck = str(deck)
        return self

    def set_threads(self, n: int) -> None:
        """n > 1: CellCapture as the reference's OpenMP backend runs it (timing only)."""
        lib().oracle_set_threads(self._h, n)

    def close(self):
        if self._h:
            lib().oracle_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def dims(self) -> dict:
        d = (ctypes.c_int64 * 17)()
        lib().oracle_dims(self._h, d)
        keys = ["X", "Y", "nz", "nzA", "nx", "ny", "nuclei", "cycle", "nn", "frontier", "captures", "solidified",
                "layers", "layer", "zlow", "zhigh", "outside"]
        return dict(zip(keys, list(d)))

    def params(self) -> dict:
        p = (ctypes.c_double * 9)()
        lib().oracle_params(self._h, p)
        return dict(zip(["A", "B", "C", "D", "deltax", "deltat", "G", "R", "dU"], list(p)))

    def field(self, which: int) -> np.ndarray:
        d = self.dims()
        n = d["X"] * d["Y"] * d["nz"]
        nact = d["X"] * d["Y"] * d["nzA"]
        shapes = {DIAGLEN: (nact, np.float32), DOCENTER: (3 * nact, np.float32), CRITDIAG: (26 * nact, np.float32),
                  NUCLEI_LOC: (d["nuclei"], np.int32), NUCLEI_TIME: (d["nuclei"], np.int32),
                  UNDERCOOLING: (n, np.float32), UCHANGE: (n, np.float32)}
        cnt, dt = shapes.get(which, (n, np.int32))
        out = np.zeros(max(cnt, 1), dtype=dt)
        if lib().oracle_get_field(self._h, which, out.ctypes.data) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return out[:cnt]

    # ---- test hooks (hand-built states, planted mistakes) ----
    def set_field(self, which: int, values: np.ndarray) -> None:
        dt = np.float32 if which in (UNDERCOOLING, UCHANGE, DIAGLEN, DOCENTER, CRITDIAG) else np.int32
        a = np.ascontiguousarray(values, dtype=dt)
        if a.size != self.field(which).size:
            raise ValueError(f"field {which}: {a.size} values for {self.field(which).size}")
        if lib().oracle_set_field(self._h, which, a.ctypes.data) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def set_orientation(self, o: int, guv9) -> None:
        g = np.ascontiguousarray(guv9, dtype=np.float32).ravel()
        if lib().oracle_set_orientation(self._h, o, _ptr(g, ctypes.c_float)) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def set_mutation(self, m: int) -> None:
        lib().oracle_set_mutation(self._h, m)

    def step(self, cycle: int) -> None:
        lib().oracle_step(self._h, cycle)

    def check(self, cycle: int, verbose: bool = False):
        c = ctypes.c_int(cycle)
        counts = (ctypes.c_uint64 * 5)()
        xs = ctypes.c_int(0)
        lib().oracle_check(self._h, ctypes.byref(c), counts, ctypes.byref(xs), int(verbose))
        return c.value, list(counts), xs.value

    def run(self, max_cycles: int = 0, verbose: bool = False):
        t = ctypes.c_double(0)
        cyc = lib().oracle_run(self._h, max_cycles, int(verbose), ctypes.byref(t))
        return cyc, t.value

    def next_layer(self) -> None:
        """DomainShiftAndResize + LayerSetup after the current layer."""
        if lib().oracle_next_layer(self._h) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def write_outputs(self) -> None:
        if lib().oracle_write_outputs(self._h, self.deck.encode()) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())


def mt64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().oracle_mt64_draws(seed, n, _ptr(out, ctypes.c_uint64))
    return out


def uniform(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().oracle_uniform_draws(seed, n, _ptr(out, ctypes.c_double))
    return out


def minstd(n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_minstd_draws(n, _ptr(out, ctypes.c_uint32))
    return out


def normal(mean: float, sd: float, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().oracle_normal_draws(mean, sd, n, _ptr(out, ctypes.c_double))
    return out


def seed_cdl(guv9) -> np.ndarray:
    g = np.ascontiguousarray(guv9, dtype=np.float32)
    out = np.zeros(26, dtype=np.float32)
    lib().oracle_seed_cdl(_ptr(g, ctypes.c_float), _ptr(out, ctypes.c_float))
    return out


def capture_octahedron(guv9, cold, p):
    g = np.ascontiguousarray(guv9, dtype=np.float32)
    centre = np.zeros(3, dtype=np.float64)
    dl = ctypes.c_float(0)
    cdl = np.zeros(26, dtype=np.float32)
    lib().oracle_capture_octahedron(_ptr(g, ctypes.c_float), float(cold[0]), float(cold[1]), float(cold[2]),
                                    float(p[0]), float(p[1]), float(p[2]), _ptr(centre, ctypes.c_double),
                                    ctypes.byref(dl), _ptr(cdl, ctypes.c_float))
    return centre, dl.value, cdl


def growth_increment(A, B, C, D, U) -> float:
    return lib().oracle_growth_increment(A, B, C, D, U)
