"""Python side of the HIP engine: a ctypes binding of include/glazier_hip.h.

`HipSimulation` is the counterpart of GpuSimulation
(glazier-cuda/src/lib.rs:749-1569): built from a model and a starting lattice,
stepped in whole Monte Carlo steps, read back into numpy arrays. The compiled
library `libglazier_hip.so` sits beside this file; without it, importing the
engine raises rather than falling back to anything else.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from .model import CModel, Model

LIB_PATH = Path(__file__).resolve().parent / "libglazier_hip.so"

GHIP_OK = 0
GHIP_ODD_LATTICE = 1
GHIP_DEVICE = 2
GHIP_UNSUPPORTED = 3
GHIP_BAD_ARG = 4


class GhipError(RuntimeError):
    """A failure the engine reported, with its status code."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_lib = None


def library() -> ctypes.CDLL:
    """The compiled engine. Raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc for gfx950)")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64
    # (same soname as ROCm's). Loading torch first lets the engine's
    # dependency resolve to the runtime already in the process; the other
    # order leaves two runtimes, and torch then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    sim = ctypes.c_void_p
    lib.ghip_create.argtypes = [P(CModel), P(ctypes.c_uint32), ctypes.c_size_t, P(ctypes.c_uint8),
                                ctypes.c_size_t, P(ctypes.c_double), P(ctypes.c_double),
                                P(ctypes.c_void_p)]
    lib.ghip_create_nematic.argtypes = [P(CModel), P(ctypes.c_uint32), ctypes.c_size_t,
                                        P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_double),
                                        P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_void_p)]
    lib.ghip_set_nematic_field.argtypes = [sim, P(ctypes.c_double), ctypes.c_size_t]
    lib.ghip_blueprint_parse.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t,
                                         P(ctypes.c_size_t)]
    lib.ghip_destroy.argtypes = [sim]
    lib.ghip_step.argtypes = [sim, ctypes.c_uint64]
    lib.ghip_mcs.argtypes = [sim]
    lib.ghip_mcs.restype = ctypes.c_uint64
    lib.ghip_n_labels.argtypes = [sim, P(ctypes.c_size_t)]
    lib.ghip_labels.argtypes = [sim, P(ctypes.c_uint32), ctypes.c_size_t]
    lib.ghip_volumes.argtypes = [sim, P(ctypes.c_uint32), ctypes.c_size_t]
    lib.ghip_surfaces.argtypes = [sim, P(ctypes.c_int64), ctypes.c_size_t]
    lib.ghip_cell_types.argtypes = [sim, P(ctypes.c_uint8), ctypes.c_size_t]
    lib.ghip_field.argtypes = [sim, ctypes.c_uint32, P(ctypes.c_double), ctypes.c_size_t]
    lib.ghip_field_f32.argtypes = [sim, ctypes.c_uint32, P(ctypes.c_float), ctypes.c_size_t]
    lib.ghip_set_field.argtypes = [sim, ctypes.c_uint32, P(ctypes.c_double), ctypes.c_size_t]
    lib.ghip_activity_f32.argtypes = [sim, P(ctypes.c_float), ctypes.c_size_t]
    lib.ghip_counters.argtypes = [sim, P(ctypes.c_uint64), P(ctypes.c_uint64)]
    lib.ghip_cell_state.argtypes = [sim, P(ctypes.c_int64), ctypes.c_size_t, P(ctypes.c_uint32)]
    lib.ghip_set_timing.argtypes = [sim, ctypes.c_int]
    lib.ghip_set_barrier_patience.argtypes = [sim, ctypes.c_uint32]
    lib.ghip_kernel_time.argtypes = [sim, ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_uint64)]
    lib.ghip_batch_create.argtypes = [P(ctypes.c_void_p), ctypes.c_size_t, P(ctypes.c_void_p)]
    lib.ghip_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.ghip_batch_step.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.ghip_batch_set_tiles.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ghip_batch_set_occupancy.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ghip_batch_describe.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.ghip_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != GHIP_OK:
        msg = library().ghip_last_error().decode()
        raise GhipError(status, msg)


def _ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(ctypes.POINTER(ctype))


class HipSimulation:
    """A run on one MI355X, seeded from a model and a lattice.

    `labels` is the flat label field (x fastest); `cell_types` one type per
    label, the medium's included (the reference keeps 1 there,
    glazier-core/src/cpu.rs:203); `fields` species-major f64 values or None
    for each species' initial value; `activity` per-site memory or None.
    """

    def __init__(self, model: Model, labels: np.ndarray, cell_types: np.ndarray,
                 fields: Optional[np.ndarray] = None, activity: Optional[np.ndarray] = None):
        lib = library()
        self.model = model
        self._c = model.to_c()
        lab = np.ascontiguousarray(labels, dtype=np.uint32).reshape(-1)
        ct = np.ascontiguousarray(cell_types, dtype=np.uint8).reshape(-1)
        self.sites = lab.size
        fptr = None
        if fields is not None:
            f = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
            fptr = _ptr(f, ctypes.c_double)
        aptr = None
        if activity is not None:
            a = np.ascontiguousarray(activity, dtype=np.float64).reshape(-1)
            aptr = _ptr(a, ctypes.c_double)
        handle = ctypes.c_void_p()
        q = model.nematic_pairs() if model.has_nematic else None
        if q is not None:
            _check(lib.ghip_create_nematic(ctypes.byref(self._c), _ptr(lab, ctypes.c_uint32), lab.size,
                                           _ptr(ct, ctypes.c_uint8), ct.size, fptr, aptr,
                                           _ptr(q, ctypes.c_double), ctypes.byref(handle)))
        else:
            _check(lib.ghip_create(ctypes.byref(self._c), _ptr(lab, ctypes.c_uint32), lab.size,
                                   _ptr(ct, ctypes.c_uint8), ct.size, fptr, aptr, ctypes.byref(handle)))
        self._h = handle

    def set_nematic_field(self, q: np.ndarray) -> None:
        """GpuSimulation::set_nematic_field (glazier-cuda/src/lib.rs:1199-1221):
        one (Q_xx, Q_xy) pair per site."""
        v = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        _check(library().ghip_set_nematic_field(self._h, _ptr(v, ctypes.c_double), v.size // 2))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(library().ghip_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping ---------------------------------------------------------
    def step(self, steps: int = 1) -> None:
        _check(library().ghip_step(self._h, int(steps)))

    @property
    def mcs(self) -> int:
        return int(library().ghip_mcs(self._h))

    # -- read-backs -------------------------------------------------------
    def n_labels(self) -> int:
        n = ctypes.c_size_t()
        _check(library().ghip_n_labels(self._h, ctypes.byref(n)))
        return int(n.value)

    def labels(self) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.uint32)
        _check(library().ghip_labels(self._h, _ptr(out, ctypes.c_uint32), out.size))
        return out

    def volumes(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.uint32)
        _check(library().ghip_volumes(self._h, _ptr(out, ctypes.c_uint32), out.size))
        return out

    def surfaces(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.int64)
        _check(library().ghip_surfaces(self._h, _ptr(out, ctypes.c_int64), out.size))
        return out

    def cell_types(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.uint8)
        _check(library().ghip_cell_types(self._h, _ptr(out, ctypes.c_uint8), out.size))
        return out

    def field(self, species: int) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float64)
        _check(library().ghip_field(self._h, species, _ptr(out, ctypes.c_double), out.size))
        return out

    def field_f32(self, species: int) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float32)
        _check(library().ghip_field_f32(self._h, species, _ptr(out, ctypes.c_float), out.size))
        return out

    def set_field(self, species: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _check(library().ghip_set_field(self._h, species, _ptr(v, ctypes.c_double), v.size))

    def activity_f32(self) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float32)
        _check(library().ghip_activity_f32(self._h, _ptr(out, ctypes.c_float), out.size))
        return out

    def counters(self):
        att, acc = ctypes.c_uint64(), ctypes.c_uint64()
        _check(library().ghip_counters(self._h, ctypes.byref(att), ctypes.byref(acc)))
        return int(att.value), int(acc.value)

    def cell_state(self) -> np.ndarray:
        """Per label: volume, surface, then the nine integer moment sums when
        the engine keeps them and the two nematic field sums (units of 2^-24)
        with the coupling; shape (n_labels, words)."""
        words = ctypes.c_uint32()
        lib = library()
        _check(lib.ghip_cell_state(self._h, None, 0, ctypes.byref(words)))
        n = self.n_labels()
        out = np.empty(n * words.value, dtype=np.int64)
        _check(lib.ghip_cell_state(self._h, _ptr(out, ctypes.c_int64), out.size, ctypes.byref(words)))
        return out.reshape(n, words.value)

    # -- timing -----------------------------------------------------------
    def set_timing(self, on) -> None:
        """True (or 1): events around every launch; 2: one event pair around
        each step call, summed under family "step"; False/0: off."""
        mode = int(on) if not isinstance(on, bool) else (1 if on else 0)
        _check(library().ghip_set_timing(self._h, mode))

    def set_barrier_patience(self, polls: int) -> None:
        """Polls a grid barrier of the tiled resident pass makes before it
        gives up (the step then raises); 0 gives up at the first barrier."""
        _check(library().ghip_set_barrier_patience(self._h, int(polls)))

    def kernel_time(self, family: str):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _check(library().ghip_kernel_time(self._h, family.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return float(ms.value), int(n.value)


class HipBatch:
    """Several HipSimulation lattices stepped together on one GPU.

    Each stage of a Monte Carlo step is one launch over every lattice, and
    after the first steps (which time the candidate tile shapes on real
    steps) the whole step is a replayed hipGraph. Every lattice computes
    exactly what its own `step` would; read it through its HipSimulation
    between batch steps. The batch holds its lattices and is closed before
    them.
    """

    def __init__(self, sims: Sequence[HipSimulation]):
        lib = library()
        self.sims = list(sims)
        handles = (ctypes.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        handle = ctypes.c_void_p()
        _check(lib.ghip_batch_create(handles, len(self.sims), ctypes.byref(handle)))
        self._h = handle

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(library().ghip_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, steps: int = 1) -> None:
        _check(library().ghip_batch_step(self._h, int(steps)))

    def set_tiles(self, candidate: int) -> None:
        """Fix every group to tile candidate 0..4, or -1 to time them again."""
        _check(library().ghip_batch_set_tiles(self._h, int(candidate)))

    def set_occupancy(self, candidate: int) -> None:
        """Fix every launch to occupancy candidate 0..3 (no cap, or at most
        4, 3 or 2 workgroups per CU) on the current tiles."""
        _check(library().ghip_batch_set_occupancy(self._h, int(candidate)))

    def describe(self) -> dict:
        import json

        buf = ctypes.create_string_buffer(1 << 16)
        _check(library().ghip_batch_describe(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())
