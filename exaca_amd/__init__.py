"""MI355X-native ExaCA cellular-automaton time step.

Python access to the C ABI declared in ``include/exaca_ca.h`` (native library
``exaca_amd/libexaca_amd.so``, built by ``exaca_amd/build.py``). The solver mirrors the
reference's driver: ``Solver(deck)`` performs the reference's initialisation for one rank and
uploads the state; ``nucleation``/``cell_capture``/``check`` are the per-step functions of
``src/CAupdate.hpp``; ``run`` is the layer loop of ``RunProgram_Reduced`` on the device.
There is no CPU fallback: a missing library is an error.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libexaca_amd.so"
DATA_DIR = PKG_DIR / "data"

FIELD_CELLTYPE = 0
FIELD_GRAINID = 1
FIELD_DIAGONAL = 2
FIELD_DOCENTER = 3
FIELD_CRITTIMESTEP = 4
FIELD_LAYERID = 5
FIELD_CELLTYPE_RAW = 6
FIELD_NUCLEI_LOC = 7
FIELD_NUCLEI_TIME = 8

# exported entry points and their signatures (include/exaca_ca.h)
_I64P = ctypes.POINTER(ctypes.c_int64)
_SIGNATURES = {
    "exaca_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "exaca_nucleation": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "exaca_cell_capture": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "exaca_check": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "exaca_run_layer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "exaca_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]),
    "exaca_query": (ctypes.c_int, [ctypes.c_void_p, _I64P]),
    "exaca_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "exaca_run_program": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
    "exaca_host_state": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_size_t, _I64P]),
    "exaca_last_error": (ctypes.c_char_p, []),
    "exaca_geometry_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "exaca_profile_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double), _I64P]),
    "exaca_copy_probe": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


class ExacaError(RuntimeError):
    """A nonzero status from the native library (message from exaca_last_error)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


def library() -> ctypes.CDLL:
    """Loads the native library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python exaca_amd/build.py`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ok(status: int) -> None:
    if status != 0:
        raise ExacaError(status, library().exaca_last_error().decode())


def _b(s) -> bytes:
    return str(s).encode()


def host_state(deck, field: int | None = None, data_dir=None):
    """Initial host state exaca_create would upload (no GPU needed). Returns dims, or
    (dims, array) when a field is requested."""
    lib = library()
    dims = (ctypes.c_int64 * 8)()
    dd = _b(data_dir) if data_dir else None
    _ok(lib.exaca_host_state(_b(deck), dd, 0, None, 0, dims))
    d = dict(zip(["X", "Y", "nz", "nzA", "nx", "ny", "nuclei", "active"], list(dims)))
    if field is None:
        return d
    n = d["nuclei"] if field in (FIELD_NUCLEI_LOC, FIELD_NUCLEI_TIME) else d["X"] * d["Y"] * d["nz"]
    out = np.zeros(max(n, 1), dtype=np.int32)
    _ok(lib.exaca_host_state(_b(deck), dd, field, out.ctypes.data, out.nbytes, dims))
    return d, out[:n]


def run_program(deck) -> None:
    """RunProgram_Reduced(0, 1, deck): the whole drop-in program on the current GPU."""
    _ok(library().exaca_run_program(0, 1, _b(deck)))


def copy_probe(nbytes: int, reps: int = 20) -> float:
    """Average device time (ms) of a plain copy of nbytes (HBM calibration)."""
    ms = ctypes.c_double(0)
    _ok(library().exaca_copy_probe(nbytes, reps, ctypes.byref(ms)))
    return ms.value


class Solver:
    """One CA domain resident on the GPU."""

    def __init__(self, deck, data_dir=None, verbose: bool = False):
        self._lib = library()
        h = ctypes.c_void_p()
        _ok(self._lib.exaca_create(_b(deck), _b(data_dir) if data_dir else None, int(verbose), ctypes.byref(h)))
        self._h = h
        q = self.query()
        self.X, self.Y, self.nz, self.nzA = q["X"], q["Y"], q["nz"], q["nzA"]
        self.ncell = self.X * self.Y * self.nz

    def close(self) -> None:
        if self._h:
            _ok(self._lib.exaca_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nucleation(self, cycle: int, layer: int = 0) -> None:
        _ok(self._lib.exaca_nucleation(self._h, cycle, layer))

    def cell_capture(self, cycle: int, layer: int = 0) -> None:
        _ok(self._lib.exaca_cell_capture(self._h, cycle, layer))

    def step(self, cycle: int, layer: int = 0) -> None:
        self.nucleation(cycle, layer)
        self.cell_capture(cycle, layer)

    def check(self, cycle: int, layer: int = 0, verbose: bool = False):
        c = ctypes.c_int(cycle)
        counts = (ctypes.c_uint64 * 5)()
        xs = ctypes.c_int(0)
        _ok(self._lib.exaca_check(self._h, ctypes.byref(c), layer, counts, ctypes.byref(xs), int(verbose)))
        return c.value, list(counts), xs.value

    def run(self, max_cycles: int = 0, layer: int = 0, verbose: bool = False):
        c = ctypes.c_int(0)
        ms = ctypes.c_double(0)
        _ok(self._lib.exaca_run_layer(self._h, layer, max_cycles, int(verbose), ctypes.byref(c), ctypes.byref(ms)))
        return c.value, ms.value

    def field(self, kind: int) -> np.ndarray:
        if kind == FIELD_CELLTYPE_RAW:
            out = np.empty(self.ncell, dtype=np.uint8)
        elif kind == FIELD_DOCENTER:
            out = np.empty(3 * self.ncell, dtype=np.float32)
        elif kind == FIELD_DIAGONAL:
            out = np.empty(self.ncell, dtype=np.float32)
        else:
            out = np.empty(self.ncell, dtype=np.int32)
        _ok(self._lib.exaca_download(self._h, kind, out.ctypes.data, out.nbytes))
        return out

    def query(self) -> dict:
        out = (ctypes.c_int64 * 16)()
        _ok(self._lib.exaca_query(self._h, out))
        keys = ["X", "Y", "nz", "nzA", "nx", "ny", "nuclei", "cycle", "nn", "frontier", "captures",
                "solidified", "waiting", "list_len", "list_cap", "device_bytes"]
        return dict(zip(keys, list(out)))

    def profile_steps(self, nsteps: int, layer: int = 0) -> dict:
        """Runs nsteps more steps with per-kernel HIP events; summed kernel times in ms."""
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int64 * 5)()
        _ok(self._lib.exaca_profile_steps(self._h, layer, nsteps, ms, cnt))
        return {"nucleation_ms": ms[0], "claim_ms": ms[1], "apply_ms": ms[2], "frontier": cnt[0],
                "captures": cnt[1], "solidified": cnt[2], "launches": cnt[3], "steps": cnt[4]}

    def geometry_probe(self, rows: np.ndarray) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 7)
        out = np.empty((rows.shape[0], 30), dtype=np.float64)
        _ok(self._lib.exaca_geometry_probe(self._h, rows.ctypes.data, out.ctypes.data, rows.shape[0]))
        return out
