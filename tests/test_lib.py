"""libsparspak_hip.so: built, loadable without a GPU, and exporting every
entry point declared in include/sparspak_hip.h (no compute calls here)."""
import ctypes
import os
import re

import pytest

from sparspak_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "sparspak_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spk(?:hip)?_[a-z0-9_]+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), "run python sparspak.jl_amd/build.py"


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    assert sorted(_lib.exported_symbols()) == _declared()


def test_loader_binds_all():
    L = _lib.lib()
    assert L.spkhip_version().startswith(b"sparspak_hip")


def test_sequence_errors_on_host():
    # phase-order errors are raised before any device call (msprs019)
    from sparspak_hip import (SparseSolver, factor, inmatrix, symbolicfactor,
                              triangularsolve)
    from sparspak_hip.gallery import laplace2d
    s = SparseSolver(laplace2d(4))
    with pytest.raises(RuntimeError, match="Ordering not done"):
        symbolicfactor(s)
    with pytest.raises(RuntimeError, match="Symbolic factor not done"):
        inmatrix(s)
    with pytest.raises(RuntimeError, match="Matrix input not done"):
        factor(s)
    with pytest.raises(RuntimeError, match="Factorization not done"):
        triangularsolve(s, [0.0] * 16)
