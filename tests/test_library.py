"""The C ABI library loads without a GPU and exports every entry point include/qm_paint.h
declares (qm_paint.h, qm_surrogate.h, qm_tiles.h); the product path refuses to run without it (no CPU fallback)."""
import ctypes
import os
import re

import pytest

from quietmap_amd import painter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for header in ("qm_paint.h", "qm_surrogate.h", "qm_tiles.h"):
        text = open(os.path.join(ROOT, "include", header)).read()
        names |= set(re.findall(r"\b(qm_[a-z_]+)\s*\(", text))
    return sorted(names)


def test_header_and_python_agree():
    assert declared() == sorted(painter.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(painter.LIBRARY):
        pytest.skip("libqmpaint.so not built")
    lib = ctypes.CDLL(painter.LIBRARY)
    for name in declared():
        assert hasattr(lib, name), name


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(painter, "LIBRARY", str(tmp_path / "missing.so"))
    monkeypatch.setattr(painter, "_lib", None)
    with pytest.raises(painter.PainterError):
        painter.library()


def test_screening_struct_matches_the_header():
    import ctypes as C

    from quietmap_amd import abi

    assert C.sizeof(abi.AircraftScreening) == 32
    assert abi.HORIZON_FLOATS == 3746 and abi.TERRAIN_HORIZON_FLOATS == 417
