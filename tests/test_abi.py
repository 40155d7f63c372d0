"""The C ABI without a GPU: the library loads, exports every function the
header declares, and refuses bad models before it touches a device."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "glazier_hip.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(ghip_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from glazier_amd import engine

    return engine.library()


def test_every_declared_function_is_exported(lib):
    names = declared()
    assert "ghip_create" in names and "ghip_batch_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_the_model_record_matches_the_header():
    from glazier_amd.model import MAX_SPECIES, MAX_TYPES, CModel

    text = HEADER.read_text()
    assert f"#define GHIP_MAX_TYPES {MAX_TYPES}" in text
    assert f"#define GHIP_MAX_SPECIES {MAX_SPECIES}" in text
    fields = [f[0] for f in CModel._fields_]
    body = text[text.index("typedef struct ghip_model"):text.index("} ghip_model;")]
    order = [f for f in fields if re.search(r"\b" + f + r"\b", body)]
    assert order == fields


def _create(lib, model, labels, types):
    from glazier_amd.engine import _ptr

    c = model.to_c()
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    ct = np.ascontiguousarray(types, dtype=np.uint8)
    h = ctypes.c_void_p()
    status = lib.ghip_create(ctypes.byref(c), _ptr(lab, ctypes.c_uint32), lab.size,
                             _ptr(ct, ctypes.c_uint8), ct.size, None, None, ctypes.byref(h))
    return status, h


def test_odd_lattices_and_unpriced_terms_are_refused(lib):
    from glazier_amd import engine
    from glazier_amd.workloads import plain_monolayer

    m, labels, types = plain_monolayer(64, 1)
    m.width = 63
    status, h = _create(lib, m, labels[:63 * 64], types)
    assert status == engine.GHIP_ODD_LATTICE and not h.value
    assert b"even" in lib.ghip_last_error()
    m, labels, types = plain_monolayer(64, 1)
    m.types[1].lambda_nematic = 1.0
    m.nematic = np.zeros((64 * 64, 2))
    status, h = _create(lib, m, labels, types)  # the field goes through ghip_create_nematic
    assert status == engine.GHIP_BAD_ARG and b"nematic" in lib.ghip_last_error()
    m, labels, types = plain_monolayer(64, 1)
    bad = types.copy()
    bad[3] = 7
    status, _ = _create(lib, m, labels, bad)
    assert status == engine.GHIP_BAD_ARG


def test_null_handles_are_refused(lib):
    from glazier_amd import engine

    out = ctypes.c_void_p()
    assert lib.ghip_batch_create(None, 0, ctypes.byref(out)) == engine.GHIP_BAD_ARG
    assert lib.ghip_step(None, 1) == engine.GHIP_BAD_ARG
    assert lib.ghip_batch_step(None, 1) == engine.GHIP_BAD_ARG
    assert lib.ghip_destroy(None) == engine.GHIP_OK
