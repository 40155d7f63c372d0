"""CPU oracles for the MI355X Cellular Potts engine - TEST INFRASTRUCTURE.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark
import this package. It binds oracle/liboracle.so (built from serial.c and
emu.c by oracle/Makefile):

  SerialSimulation  the restatement of glazier-core's serial f64 engine
                    (glazier-core/src/cpu.rs): the distributional reference
                    and the timed CPU baseline;
  EmuSimulation     the CPU emulation of the device checkerboard schedule:
                    the bit-exact reference for the HIP kernels;
  replay_forward / replay_gradient
                    the f64 replay of the field step over a recorded type
                    trajectory and its adjoint (adjoint.c): the reference
                    for the device gradient.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

from glazier_amd.model import CFieldParams, CModel, FieldParams, Model

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    v = ctypes.c_void_p
    u32, u64, i64, f64, f32, sz = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64,
                                   ctypes.c_double, ctypes.c_float, ctypes.c_size_t)
    sigs = {
        "oracle_splitmix64_next": ([P(u64)], u64),
        "oracle_xoshiro_seed": ([u64, P(u64)], None),
        "oracle_xoshiro_next": ([P(u64)], u64),
        "oracle_xoshiro_f64": ([P(u64)], f64),
        "oracle_xoshiro_below": ([P(u64), u64], u64),
        "serial_create": ([P(CModel), P(u32), P(ctypes.c_uint8), sz, P(f64)], v),
        "serial_destroy": ([v], None),
        "serial_step": ([v, u64], None),
        "serial_attempt": ([v], ctypes.c_int),
        "serial_delta_energy": ([v, sz, u32], f64),
        "serial_energy": ([v], f64),
        "serial_chemotaxis_work": ([v, u32, sz, sz], f64),
        "serial_move_work": ([v, u32, u32, sz, sz], f64),
        "serial_divide_and_die": ([v, P(sz), P(sz)], None),
        "serial_n_labels": ([v], sz),
        "serial_mcs": ([v], u64),
        "serial_counters": ([v, P(u64), P(u64)], None),
        "serial_labels": ([v, P(u32)], None),
        "serial_set_labels": ([v, P(u32)], None),
        "serial_volumes": ([v, P(u32)], None),
        "serial_surfaces": ([v, P(i64)], None),
        "serial_cell_types": ([v, P(ctypes.c_uint8)], None),
        "serial_recount_volumes": ([v, P(u32)], None),
        "serial_recount_surfaces": ([v, P(i64)], None),
        "serial_recount": ([v], None),
        "serial_lengths": ([v, P(f64)], None),
        "serial_recounted_lengths": ([v, P(f64)], None),
        "serial_field": ([v, u32, P(f64)], None),
        "serial_set_field": ([v, u32, P(f64)], None),
        "serial_activity": ([v, P(f64)], None),
        "serial_set_activity": ([v, P(f64)], None),
        "serial_fields_step": ([v], None),
        "serial_locally_connected": ([v, sz, u32], ctypes.c_int),
        "serial_substeps": ([u32, P(CModel)], ctypes.c_int),
        "serial_clone": ([v], v),
        "serial_apply_copy": ([v, sz, u32], None),
        "serial_set_nematic": ([v, P(f64)], None),
        "serial_decay_activity": ([v], None),
        "serial_nematic_sums": ([v, P(f64)], None),
        "serial_recounted_nematic": ([v, P(f64)], None),
        "serial_recounted_anisotropy": ([v, P(f64)], None),
        "emu_set_nematic": ([v, P(f64)], None),
        "emu_nematic_fixed": ([f64], ctypes.c_int32),
        "oracle_moments_length": ([P(f64), sz], f64),
        "oracle_principal_axis": ([P(f64), P(f64), P(f64)], None),
        "oracle_hash32": ([u32], u32),
        "oracle_pass_key": ([u64, u32, u64, u32], u32),
        "oracle_site_pick": ([u32, u32], u32),
        "oracle_site_draw": ([u32, u32], u32),
        "emu_create": ([P(CModel), P(u32), P(ctypes.c_uint8), sz, P(f64), P(f64)], v),
        "emu_destroy": ([v], None),
        "emu_step": ([v, u64], None),
        "emu_n_labels": ([v], sz),
        "emu_mcs": ([v], u64),
        "emu_counters": ([v, P(u64), P(u64)], None),
        "emu_labels": ([v, P(u32)], None),
        "emu_volumes": ([v, P(u32)], None),
        "emu_surfaces": ([v, P(i64)], None),
        "emu_cell_types": ([v, P(ctypes.c_uint8)], None),
        "emu_cell_state": ([v, P(i64)], u32),
        "emu_field_f32": ([v, u32, P(f32)], None),
        "emu_set_field": ([v, u32, P(f64)], None),
        "emu_activity_f32": ([v, P(f32)], None),
        "emu_delta_energy": ([v, sz, sz], f64),
        "emu_length": ([P(i64), i64, ctypes.c_int], f64),
        "emu_fields_step": ([v], None),
        "emu_one_piece": ([u32, ctypes.c_int], ctypes.c_int),
        "emu_activity_mean": ([v, sz, u32], f32),
        "serial_activity_mean": ([v, sz, u32], f64),
        "oracle_exp2_f32": ([f32], f32),
        "oracle_log2_f32": ([f32], f32),
        "oracle_ring_table": ([ctypes.c_int, u32, u32, P(ctypes.c_uint8)], None),
        "adj_forward": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, u32, P(ctypes.c_int),
                         P(CFieldParams), P(ctypes.c_uint8), u64, P(f64), P(f64)], None),
        "adj_gradient": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, u32, P(ctypes.c_int),
                          P(CFieldParams), P(ctypes.c_uint8), u64, P(f64), P(f64),
                          P(CFieldParams), P(f64)], None),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Xoshiro:
    """xoshiro256++ seeded through SplitMix64 (glazier-core/src/rng.rs)."""

    def __init__(self, seed: int):
        self.s = (ctypes.c_uint64 * 4)()
        library().oracle_xoshiro_seed(seed, self.s)

    def next_u64(self) -> int:
        return int(library().oracle_xoshiro_next(self.s))

    def next_f64(self) -> float:
        return float(library().oracle_xoshiro_f64(self.s))

    def below(self, n: int) -> int:
        return int(library().oracle_xoshiro_below(self.s, n))


class _Base:
    _prefix = ""

    def _f(self, name):
        return getattr(library(), self._prefix + name)

    def labels(self) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.uint32)
        self._f("labels")(self._h, _p(out, ctypes.c_uint32))
        return out

    def n_labels(self) -> int:
        return int(self._f("n_labels")(self._h))

    def volumes(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.uint32)
        self._f("volumes")(self._h, _p(out, ctypes.c_uint32))
        return out

    def surfaces(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.int64)
        self._f("surfaces")(self._h, _p(out, ctypes.c_int64))
        return out

    def cell_types(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.uint8)
        self._f("cell_types")(self._h, _p(out, ctypes.c_uint8))
        return out

    def counters(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self._f("counters")(self._h, ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    @property
    def mcs(self) -> int:
        return int(self._f("mcs")(self._h))

    def step(self, steps: int = 1) -> None:
        self._f("step")(self._h, int(steps))

    def close(self):
        if getattr(self, "_h", None):
            self._f("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SerialSimulation(_Base):
    """The serial f64 engine (glazier-core/src/cpu.rs) from a model, a label
    field and one type per label."""

    _prefix = "serial_"

    def __init__(self, model: Model, labels: np.ndarray, cell_types: np.ndarray,
                 fields: Optional[np.ndarray] = None):
        self.model = model
        self._c = model.to_c()
        lab = np.ascontiguousarray(labels, dtype=np.uint32).reshape(-1)
        ct = np.ascontiguousarray(cell_types, dtype=np.uint8).reshape(-1)
        self.sites = lab.size
        fp = None
        if fields is not None:
            self._fields = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
            fp = _p(self._fields, ctypes.c_double)
        self._h = library().serial_create(ctypes.byref(self._c), _p(lab, ctypes.c_uint32),
                                          _p(ct, ctypes.c_uint8), ct.size, fp)
        q = model.nematic_pairs()
        if q is not None:
            library().serial_set_nematic(self._h, _p(q, ctypes.c_double))

    def clone(self) -> "SerialSimulation":
        c = SerialSimulation.__new__(SerialSimulation)
        c.model, c._c, c.sites = self.model, self._c, self.sites
        c._h = library().serial_clone(self._h)
        return c

    def apply_copy(self, target: int, new: int) -> None:
        library().serial_apply_copy(self._h, target, new)

    def attempt(self) -> bool:
        return bool(library().serial_attempt(self._h))

    def delta_energy(self, target: int, new: int) -> float:
        return float(library().serial_delta_energy(self._h, target, new))

    def energy(self) -> float:
        return float(library().serial_energy(self._h))

    def chemotaxis_work(self, new: int, target: int, source: int) -> float:
        return float(library().serial_chemotaxis_work(self._h, new, target, source))

    def move_work(self, new: int, old: int, target: int, source: int) -> float:
        return float(library().serial_move_work(self._h, new, old, target, source))

    def divide_and_die(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        library().serial_divide_and_die(self._h, ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    def set_labels(self, labels: np.ndarray, recount: bool = True) -> None:
        lab = np.ascontiguousarray(labels, dtype=np.uint32).reshape(-1)
        library().serial_set_labels(self._h, _p(lab, ctypes.c_uint32))
        if recount:
            library().serial_recount(self._h)

    def recounted_volumes(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.uint32)
        library().serial_recount_volumes(self._h, _p(out, ctypes.c_uint32))
        return out

    def recounted_surfaces(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.int64)
        library().serial_recount_surfaces(self._h, _p(out, ctypes.c_int64))
        return out

    def lengths(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.float64)
        library().serial_lengths(self._h, _p(out, ctypes.c_double))
        return out

    def recounted_lengths(self) -> np.ndarray:
        out = np.empty(self.n_labels(), dtype=np.float64)
        library().serial_recounted_lengths(self._h, _p(out, ctypes.c_double))
        return out

    def field(self, species: int) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float64)
        library().serial_field(self._h, species, _p(out, ctypes.c_double))
        return out

    def set_field(self, species: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        library().serial_set_field(self._h, species, _p(v, ctypes.c_double))

    def activity(self) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float64)
        library().serial_activity(self._h, _p(out, ctypes.c_double))
        return out

    def set_activity(self, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        library().serial_set_activity(self._h, _p(v, ctypes.c_double))

    def fields_step(self) -> None:
        library().serial_fields_step(self._h)

    def locally_connected(self, target: int, label: int) -> bool:
        return bool(library().serial_locally_connected(self._h, target, label))

    def activity_mean(self, site: int, label: int) -> float:
        """Activity::neighbourhood_mean (motility.rs:56-78) at `site` over `label`."""
        return float(library().serial_activity_mean(self._h, site, label))

    def decay_activity(self) -> None:
        """One Activity::decay (motility.rs:38-45)."""
        library().serial_decay_activity(self._h)

    def set_nematic_field(self, q: np.ndarray) -> None:
        """Model::nematic replaced (2 values per site); the sums are recounted."""
        v = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if v.size != 2 * self.sites:
            raise ValueError(f"the nematic field has {v.size // 2} sites, the lattice {self.sites}")
        library().serial_set_nematic(self._h, _p(v, ctypes.c_double))

    def nematic_sums(self) -> np.ndarray:
        out = np.empty((self.n_labels(), 2), dtype=np.float64)
        library().serial_nematic_sums(self._h, _p(out, ctypes.c_double))
        return out

    def recounted_nematic(self) -> np.ndarray:
        out = np.empty((self.n_labels(), 2), dtype=np.float64)
        library().serial_recounted_nematic(self._h, _p(out, ctypes.c_double))
        return out

    def recounted_anisotropy(self) -> np.ndarray:
        """Each label's (a0, a1) (moments.rs:131-141) from recounted moments."""
        out = np.empty((self.n_labels(), 2), dtype=np.float64)
        library().serial_recounted_anisotropy(self._h, _p(out, ctypes.c_double))
        return out


class EmuSimulation(_Base):
    """The CPU emulation of the device schedule (oracle/emu.c)."""

    _prefix = "emu_"

    def __init__(self, model: Model, labels: np.ndarray, cell_types: np.ndarray,
                 fields: Optional[np.ndarray] = None, activity: Optional[np.ndarray] = None):
        self.model = model
        self._c = model.to_c()
        lab = np.ascontiguousarray(labels, dtype=np.uint32).reshape(-1)
        ct = np.ascontiguousarray(cell_types, dtype=np.uint8).reshape(-1)
        self.sites = lab.size
        fp = ap = None
        if fields is not None:
            f = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
            fp = _p(f, ctypes.c_double)
        if activity is not None:
            a = np.ascontiguousarray(activity, dtype=np.float64).reshape(-1)
            ap = _p(a, ctypes.c_double)
        self._h = library().emu_create(ctypes.byref(self._c), _p(lab, ctypes.c_uint32),
                                       _p(ct, ctypes.c_uint8), ct.size, fp, ap)
        q = model.nematic_pairs()
        if q is not None:
            library().emu_set_nematic(self._h, _p(q, ctypes.c_double))

    def set_nematic_field(self, q: np.ndarray) -> None:
        v = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if v.size != 2 * self.sites:
            raise ValueError(f"the nematic field has {v.size // 2} sites, the lattice {self.sites}")
        library().emu_set_nematic(self._h, _p(v, ctypes.c_double))

    def cell_state(self) -> np.ndarray:
        lib = library()
        words = int(lib.emu_cell_state(self._h, None))
        out = np.empty(self.n_labels() * words, dtype=np.int64)
        lib.emu_cell_state(self._h, _p(out, ctypes.c_int64))
        return out.reshape(-1, words)

    def field_f32(self, species: int) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float32)
        library().emu_field_f32(self._h, species, _p(out, ctypes.c_float))
        return out

    def set_field(self, species: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        library().emu_set_field(self._h, species, _p(v, ctypes.c_double))

    def activity_f32(self) -> np.ndarray:
        out = np.empty(self.sites, dtype=np.float32)
        library().emu_activity_f32(self._h, _p(out, ctypes.c_float))
        return out

    def delta_energy(self, target: int, source: int) -> float:
        return float(library().emu_delta_energy(self._h, target, source))

    def activity_mean(self, site: int, label: int) -> float:
        """The pricing's f32 geometric mean of the activity over `label` around
        `site`, on the current lattice."""
        return float(library().emu_activity_mean(self._h, site, label))

    def fields_step(self) -> None:
        library().emu_fields_step(self._h)


def hash32(x: int) -> int:
    return int(library().oracle_hash32(x))


def one_piece(mask: int, ring_size: int) -> bool:
    return bool(library().emu_one_piece(mask, ring_size))


def _substeps(model: Model):
    return np.array([library().serial_substeps(k, ctypes.byref(model.to_c()))
                     for k in range(len(model.species))], dtype=np.int32)


def replay_forward(model: Model, params: FieldParams, types: np.ndarray,
                   u0: np.ndarray) -> np.ndarray:
    """f64 fields after replaying `types` (steps x n_sites, 0 = medium) from
    `u0` (n_species x n_sites) with `params` (adjoint.c, adj_forward)."""
    types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1, model.sites)
    u0 = np.ascontiguousarray(u0, dtype=np.float64).reshape(len(model.species), -1)
    sub = _substeps(model)
    out = np.empty_like(u0)
    c = params.to_c()
    library().adj_forward(model.width, model.height, model.depth, len(model.species),
                          _p(sub, ctypes.c_int), ctypes.byref(c), _p(types, ctypes.c_uint8),
                          types.shape[0], _p(u0, ctypes.c_double), _p(out, ctypes.c_double))
    return out


def replay_gradient(model: Model, params: FieldParams, types: np.ndarray, u0: np.ndarray,
                    grad_out: np.ndarray):
    """(dL/dparams, dL/du0) of a loss with dL/du_final = grad_out, in f64
    (adjoint.c, adj_gradient)."""
    types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1, model.sites)
    u0 = np.ascontiguousarray(u0, dtype=np.float64).reshape(len(model.species), -1)
    g = np.ascontiguousarray(grad_out, dtype=np.float64).reshape(u0.shape)
    sub = _substeps(model)
    gu0 = np.empty_like(u0)
    c, gc = params.to_c(), CFieldParams()
    library().adj_gradient(model.width, model.height, model.depth, len(model.species),
                           _p(sub, ctypes.c_int), ctypes.byref(c), _p(types, ctypes.c_uint8),
                           types.shape[0], _p(u0, ctypes.c_double), _p(g, ctypes.c_double),
                           ctypes.byref(gc), _p(gu0, ctypes.c_double))
    return FieldParams.from_c(gc, params.n_types, params.n_species), gu0
