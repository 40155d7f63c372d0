"""The session surface of the reference's Python module, on the HIP engine.

`glazier._native` (glazier-py/src/lib.rs) offers `parse`, `run` and
`run_volumes` over a Blueprint description, and `GpuSession`, a run kept on
the device between calls for coupling to another solver. The same names and
argument meanings are here:

* `parse(text)` is this engine's reader (the C++ reader of the `glazier`
  binary, through `ghip_blueprint_parse`), so it can be compared against the
  Python reader in `glazier_amd.blueprint` as the reference compares its two;
* `run(text, steps=None)` and `run_volumes(text, steps=None)` build the
  simulation the description states and step it. The reference steps its
  serial engine there (lib.rs:152-178); these step the HIP engine, the path
  this package accelerates;
* `GpuSession(text, base=".")` keeps one run on the device: `step`,
  `labels`, `set_nematic_field` and the `mcs` getter (lib.rs:184-241).

Errors in a description raise ValueError, as the reference maps them to
PyValueError.
"""

from __future__ import annotations

import ctypes
import json
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .blueprint import Blueprint
from .engine import GHIP_OK, GhipError, HipSimulation, library


def parse(text: str) -> dict:
    """What this engine read from a description (lib.rs:18-150)."""
    lib = library()
    raw = text.encode()
    needed = ctypes.c_size_t()
    if lib.ghip_blueprint_parse(raw, None, 0, ctypes.byref(needed)) != GHIP_OK:
        raise ValueError(lib.ghip_last_error().decode())
    buf = ctypes.create_string_buffer(needed.value)
    if lib.ghip_blueprint_parse(raw, buf, needed.value, ctypes.byref(needed)) != GHIP_OK:
        raise ValueError(lib.ghip_last_error().decode())
    return json.loads(buf.value.decode())


def _simulation(text: str, base: str) -> Tuple[Blueprint, HipSimulation]:
    parse(text)  # the engine's own checks first, with its messages
    bp = Blueprint.from_json(text, base=base)
    model, labels, types = bp.simulation()
    try:
        return bp, HipSimulation(model, labels, types)
    except GhipError as e:
        raise ValueError(str(e)) from e


def run(text: str, steps: Optional[int] = None) -> Tuple[List[int], int, int, int]:
    """Run a description and return the lattice flat (x fastest, z slowest)
    with its width, height and depth (lib.rs:152-164)."""
    bp, sim = _simulation(text, ".")
    try:
        sim.step(bp.steps if steps is None else int(steps))
        return sim.labels().tolist(), bp.width, bp.height, bp.depth
    finally:
        sim.close()


def run_volumes(text: str, steps: Optional[int] = None) -> List[int]:
    """Per-cell volumes after a run, index 0 the medium (lib.rs:166-178)."""
    bp, sim = _simulation(text, ".")
    try:
        sim.step(bp.steps if steps is None else int(steps))
        return sim.volumes().tolist()
    finally:
        sim.close()


class GpuSession:
    """A run kept on the device between calls (lib.rs:184-241). Relative
    `.npy` paths in the description resolve against `base`."""

    def __init__(self, text: str, base: str = "."):
        bp, self._sim = _simulation(text, base)
        self._sites = bp.width * bp.height * bp.depth

    def step(self, steps: int) -> None:
        """Run `steps` Monte Carlo steps."""
        try:
            self._sim.step(int(steps))
        except GhipError as e:
            raise ValueError(str(e)) from e

    def labels(self) -> List[int]:
        """The lattice, flat with x fastest."""
        return self._sim.labels().tolist()

    def set_nematic_field(self, flat: Sequence[float]) -> None:
        """Replace the field from a flat list of 2 * sites numbers, (Q_xx, Q_xy)
        per site."""
        if len(flat) != 2 * self._sites:
            raise ValueError(f"expected {2 * self._sites} numbers, got {len(flat)}")
        try:
            self._sim.set_nematic_field(np.asarray(flat, dtype=np.float64))
        except GhipError as e:
            raise ValueError(str(e)) from e

    @property
    def mcs(self) -> int:
        """Monte Carlo steps completed."""
        return self._sim.mcs

    def close(self) -> None:
        self._sim.close()
