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
# The same kernels compiled with floating-point contraction allowed: results are judged by grain
# statistics (count, mean volume, misorientation histogram), not bit for bit.
FAST_LIB_PATH = PKG_DIR / "libexaca_amd_fast.so"
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
FIELD_MELTED = 9
FIELD_CENTRE_F64 = 10

# exported entry points and their signatures (include/exaca_ca.h)
_I64P = ctypes.POINTER(ctypes.c_int64)
_SIGNATURES = {
    "exaca_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "exaca_nucleation": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "exaca_nucleation_nn": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]),
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
    "exaca_profile_kernels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double), _I64P]),
    "exaca_create_seeded": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "exaca_prepare": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "exaca_prepared_info": (ctypes.c_int, [ctypes.c_void_p, _I64P]),
    "exaca_prepared_free": (ctypes.c_int, [ctypes.c_void_p]),
    "exaca_ensemble_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "exaca_ensemble_admit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "exaca_ensemble_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "exaca_ensemble_advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "exaca_ensemble_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32)]),
    "exaca_ensemble_set_blocks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "exaca_ensemble_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "exaca_ensemble_query": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _I64P]),
    "exaca_ensemble_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_size_t]),
    "exaca_ensemble_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "exaca_grain_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), _I64P,
                                         ctypes.POINTER(ctypes.c_double), _I64P]),
    "exaca_grain_stats_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), _I64P,
                                                 ctypes.POINTER(ctypes.c_double), _I64P]),
    "exaca_next_layer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
}

_libs = {}


class ExacaError(RuntimeError):
    """A nonzero status from the native library (message from exaca_last_error)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


def library(fast: bool = False) -> ctypes.CDLL:
    """Loads the native library (or its fast build) once; raises if it has not been built."""
    if fast not in _libs:
        path = FAST_LIB_PATH if fast else LIB_PATH
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python exaca_amd/build.py`")
        lib = ctypes.CDLL(str(path))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)  # every entry point of include/exaca_ca.h must be exported
            fn.restype = res
            fn.argtypes = args
        _libs[fast] = lib
    return _libs[fast]


def _ok(status: int, lib: ctypes.CDLL | None = None) -> None:
    if status != 0:
        raise ExacaError(status, (lib or library()).exaca_last_error().decode())


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


class Prepared:
    """A type-C deck initialised on the host (any thread, no GPU), ready for an Ensemble."""

    def __init__(self, deck, seed: int = 0, data_dir=None):
        lib = library()
        h = ctypes.c_void_p()
        _ok(lib.exaca_prepare(_b(deck), _b(data_dir) if data_dir else None, seed, ctypes.byref(h)))
        self._h = h
        info = (ctypes.c_int64 * 4)()
        _ok(lib.exaca_prepared_info(h, info))
        self.cells, self.device_bytes, self.nuclei, self.liquid0 = list(info)

    def close(self):
        if self._h:
            _ok(library().exaca_prepared_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Many domains resident on one GPU, stepped together in batched launches."""

    def _ok(self, status: int) -> None:
        _ok(status, self._lib)

    def __init__(self, slots: int, arena_bytes: int):
        self._lib = library()
        h = ctypes.c_void_p()
        self._ok(self._lib.exaca_ensemble_create(slots, arena_bytes, ctypes.byref(h)))
        self._h = h
        self.slots = slots

    def close(self):
        if self._h:
            self._ok(self._lib.exaca_ensemble_destroy(self._h))
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

    def admit(self, prepared: Prepared) -> int | None:
        """Slot index, or None when no slot or memory is free."""
        s = ctypes.c_int(-1)
        st = self._lib.exaca_ensemble_admit(self._h, prepared._h, ctypes.byref(s))
        if st == 3:
            return None
        _ok(st)
        return s.value

    def release(self, slot: int) -> None:
        self._ok(self._lib.exaca_ensemble_release(self._h, slot))

    def advance(self, steps: int) -> None:
        self._ok(self._lib.exaca_ensemble_advance(self._h, steps))

    def poll(self) -> np.ndarray:
        """Final cycle per slot (0 while running, -1 when empty); see poll_errors for failures."""
        return self.poll_errors()[0]

    def poll_errors(self) -> tuple[np.ndarray, np.ndarray]:
        """(final cycles, device error flags) per slot. A slot with a nonzero flag was stopped by
        the device at the cycle given and holds no valid result."""
        out = np.zeros(self.slots, dtype=np.int32)
        err = np.zeros(self.slots, dtype=np.int32)
        p = ctypes.POINTER(ctypes.c_int32)
        self._ok(self._lib.exaca_ensemble_poll(self._h, out.ctypes.data_as(p), err.ctypes.data_as(p)))
        return out, err

    def set_blocks(self, blocks: int) -> None:
        self._ok(self._lib.exaca_ensemble_set_blocks(self._h, blocks))

    def autotune(self, chunks_per_candidate: int = 2) -> int:
        c = ctypes.c_int(0)
        self._ok(self._lib.exaca_ensemble_autotune(self._h, chunks_per_candidate, ctypes.byref(c)))
        return c.value

    def query(self, slot: int) -> dict:
        out = (ctypes.c_int64 * 8)()
        self._ok(self._lib.exaca_ensemble_query(self._h, slot, out))
        return dict(zip(["final_cycle", "nn", "frontier", "captures", "solidified", "error", "free_bytes",
                         "steps"], list(out)))

    def field(self, slot: int, kind: int, ncell: int) -> np.ndarray:
        out = np.empty(ncell, dtype=np.int32)
        self._ok(self._lib.exaca_ensemble_download(self._h, slot, kind, out.ctypes.data, out.nbytes))
        return out


class Solver:
    """One CA domain resident on the GPU. fast=True runs the fast build (contraction allowed)."""

    def _ok(self, status: int) -> None:
        _ok(status, self._lib)

    def __init__(self, deck, data_dir=None, verbose: bool = False, seed: int = 0, fast: bool = False):
        self._lib = library(fast)
        h = ctypes.c_void_p()
        self._ok(self._lib.exaca_create_seeded(_b(deck), _b(data_dir) if data_dir else None, seed, int(verbose),
                                          ctypes.byref(h)))
        self._h = h
        q = self.query()
        self.X, self.Y, self.nz, self.nzA = q["X"], q["Y"], q["nz"], q["nzA"]
        self.ncell = self.X * self.Y * self.nz

    def close(self) -> None:
        if self._h:
            self._ok(self._lib.exaca_destroy(self._h))
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
        self._ok(self._lib.exaca_nucleation(self._h, cycle, layer))

    def nucleation_nn(self, cycle: int, layer: int = 0) -> int:
        """Nucleation returning the reference's nn (the run's nucleation count so far)."""
        nn = ctypes.c_int32(0)
        self._ok(self._lib.exaca_nucleation_nn(self._h, cycle, layer, ctypes.byref(nn)))
        return nn.value

    def cell_capture(self, cycle: int, layer: int = 0) -> None:
        self._ok(self._lib.exaca_cell_capture(self._h, cycle, layer))

    def step(self, cycle: int, layer: int = 0) -> None:
        self.nucleation(cycle, layer)
        self.cell_capture(cycle, layer)

    def check(self, cycle: int, layer: int = 0, verbose: bool = False):
        c = ctypes.c_int(cycle)
        counts = (ctypes.c_uint64 * 5)()
        xs = ctypes.c_int(0)
        self._ok(self._lib.exaca_check(self._h, ctypes.byref(c), layer, counts, ctypes.byref(xs), int(verbose)))
        return c.value, list(counts), xs.value

    def run(self, max_cycles: int = 0, layer: int | None = None, verbose: bool = False):
        """The layer loop of RunProgram_Reduced on the device (exaca_run_layer): with max_cycles 0
        the current layer runs to completion and every later layer follows (exaca_next_layer in
        between); otherwise the current layer stops after max_cycles. Returns the final cycle (of
        the last layer run) and the device time in ms."""
        q = self.query()
        layer = q["layer"] if layer is None else layer
        total = 0.0
        while True:
            c = ctypes.c_int(0)
            ms = ctypes.c_double(0)
            self._ok(self._lib.exaca_run_layer(self._h, layer, max_cycles, int(verbose), ctypes.byref(c),
                                               ctypes.byref(ms)))
            total += ms.value
            if max_cycles > 0 or layer + 1 >= q["layers"]:
                return c.value, total
            self.next_layer(layer, verbose)
            layer += 1

    def next_layer(self, layer: int, verbose: bool = False) -> None:
        """DomainShiftAndResize + LayerSetup after `layer` (exaca_next_layer)."""
        self._ok(self._lib.exaca_next_layer(self._h, layer, int(verbose)))

    def field(self, kind: int) -> np.ndarray:
        if kind in (FIELD_NUCLEI_LOC, FIELD_NUCLEI_TIME):
            out = np.zeros(max(self.query()["nuclei"], 1), dtype=np.int32)
            self._ok(self._lib.exaca_download(self._h, kind, out.ctypes.data, out.nbytes))
            return out[:self.query()["nuclei"]]
        if kind == FIELD_CELLTYPE_RAW:
            out = np.empty(self.ncell, dtype=np.uint8)
        elif kind == FIELD_DOCENTER:
            out = np.empty(3 * self.ncell, dtype=np.float32)
        elif kind == FIELD_CENTRE_F64:
            out = np.empty(3 * self.ncell, dtype=np.float64)
        elif kind == FIELD_DIAGONAL:
            out = np.empty(self.ncell, dtype=np.float32)
        else:
            out = np.empty(self.ncell, dtype=np.int32)
        self._ok(self._lib.exaca_download(self._h, kind, out.ctypes.data, out.nbytes))
        return out

    def front_state(self) -> dict:
        """The whole state of the last step, as the reference holds it: cell types with the
        cell-centred flag (uint8), GrainID, DiagonalLength, octahedron centres in double, and the
        nucleation sites."""
        return {"ct_raw": self.field(FIELD_CELLTYPE_RAW), "gid": self.field(FIELD_GRAINID),
                "dl": self.field(FIELD_DIAGONAL), "centre": self.field(FIELD_CENTRE_F64),
                "nloc": self.field(FIELD_NUCLEI_LOC), "ntime": self.field(FIELD_NUCLEI_TIME)}

    def grain_stats(self, region=None, noalloc: bool = False) -> dict:
        """Grain statistics of a region (XMin, XMax, YMin, YMax, ZMin, ZMax, inclusive, in the
        coordinates of the reference's vtk files; default: every cell inside the walls), computed on
        the GPU from the current GrainID field (exaca_grain_stats, or exaca_grain_stats_noalloc,
        which allocates nothing)."""
        q = self.query()
        if region is None:
            region = (1, q["nx"] - 2, 1, q["ny"] - 2, 1, q["nz"] - 2)
        reg = (ctypes.c_int32 * 6)(*region)
        counts = (ctypes.c_int64 * 5)()
        values = (ctypes.c_double * 4)()
        hist = (ctypes.c_int64 * 64)()
        fn = self._lib.exaca_grain_stats_noalloc if noalloc else self._lib.exaca_grain_stats
        self._ok(fn(self._h, reg, counts, values, hist))
        return {"cells": counts[0], "melted": counts[1], "grains": counts[2], "nucleated_cells": counts[3],
                "melted_top": counts[4], "mean_misorientation": values[0], "mean_misorientation_top": values[1],
                "mean_grain_volume": values[2], "nucleated_fraction": values[3], "histogram": list(hist)}

    def query(self) -> dict:
        out = (ctypes.c_int64 * 20)()
        self._ok(self._lib.exaca_query(self._h, out))
        keys = ["X", "Y", "nz", "nzA", "nx", "ny", "nuclei", "cycle", "nn", "frontier", "captures",
                "solidified", "waiting", "list_len", "list_cap", "device_bytes", "evaluated", "masks", "layers",
                "layer"]
        return dict(zip(keys, list(out)))

    def profile_steps(self, nsteps: int, layer: int = 0) -> dict:
        """Runs nsteps more steps with per-kernel HIP events; summed kernel times in ms."""
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int64 * 5)()
        self._ok(self._lib.exaca_profile_steps(self._h, layer, nsteps, ms, cnt))
        return {"nucleation_ms": ms[0], "claim_ms": ms[1], "apply_ms": ms[2], "frontier": cnt[0],
                "captures": cnt[1], "solidified": cnt[2], "launches": cnt[3], "steps": cnt[4]}

    def profile_kernels(self, nsteps: int, layer: int = 0) -> dict:
        """Runs nsteps more steps with an event pair around every kernel; summed times in ms."""
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_int64 * 5)()
        self._ok(self._lib.exaca_profile_kernels(self._h, layer, nsteps, ms, cnt))
        names = ["k_nucleate", "k_claim", "k_win", "k_apply", "k_settle", "k_rebuild"]
        return {"ms": dict(zip(names, list(ms))), "frontier": cnt[0], "captures": cnt[1], "solidified": cnt[2],
                "steps": cnt[4]}

    def geometry_probe(self, rows: np.ndarray) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 7)
        out = np.empty((rows.shape[0], 30), dtype=np.float64)
        self._ok(self._lib.exaca_geometry_probe(self._h, rows.ctypes.data, out.ctypes.data, rows.shape[0]))
        return out
