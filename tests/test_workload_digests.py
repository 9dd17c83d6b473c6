"""The synthetic benchmark workloads are pinned: every generator's output hashes to the
SHA-256 committed in tests/golden/workload_sha256.json (written by
tests/golden/make_workload_sha256.py).  A change of numpy's random streams or of a
generator fails here instead of silently changing the measured workloads."""
import importlib.util
import json

import pytest

from conftest import GOLDEN

_spec = importlib.util.spec_from_file_location("make_workload_sha256", GOLDEN / "make_workload_sha256.py")
_mk = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_mk)
_WORKLOADS = _mk.workloads()
_COMMITTED = json.loads((GOLDEN / "workload_sha256.json").read_text())["digests"]


def test_every_workload_has_a_digest():
    assert {w[0] for w in _WORKLOADS} == set(_COMMITTED)


@pytest.mark.parametrize("name,desc,fn", _WORKLOADS, ids=[w[0] for w in _WORKLOADS])
def test_workload_digest(name, desc, fn):
    assert fn() == _COMMITTED[name]["sha256"], f"{name} ({desc}) changed"


def test_digest_detects_a_one_ulp_change():
    import numpy as np

    a = np.linspace(-1, 1, 1000, dtype=np.float32)
    b = a.copy()
    b[500] = np.nextafter(b[500], np.float32(2))
    assert _mk.digest(a) != _mk.digest(b)
    assert _mk.digest(a) != _mk.digest(a.astype(np.float64))  # the dtype is part of the digest
