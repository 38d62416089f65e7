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
        names |= set(re.findall(r"\b(qm_[a-z0-9_]+)\s*\(", text))
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


def test_flight_records_match_the_header(tmp_path):
    """abi's flight record dtypes have the C structs' sizes and field offsets (qm_aircraft.h)."""
    import shutil
    import subprocess

    from quietmap_amd import abi

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    header = os.path.join(os.path.dirname(painter.HERE), "quietmap_amd", "csrc", "qm_aircraft.h")
    checks = {
        "qm_flight_box": (abi.FLIGHT_BOX_DTYPE, ["east_m", "passing", "piece_length_m", "levels_db", "tail_levels_db",
                                                 "lg_scaled_distance", "installation_shares", "screened", "silent"]),
        "qm_flight_point": (abi.FLIGHT_POINT_DTYPE, ["x", "altitude_m", "horizons", "impedance_gain"]),
        "qm_flight_segment": (abi.FLIGHT_SEGMENT_DTYPE, ["within_m", "point", "sum", "first", "count", "weight",
                                                         "coarse_m"]),
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{header}"', "int main(void) {"]
    for struct, (_, fields) in checks.items():
        lines.append(f'printf("%zu\\n", sizeof({struct}));')
        for f in fields:
            lines.append(f'printf("%zu\\n", offsetof({struct}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    values = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = []
    for struct, (dtype, fields) in checks.items():
        want.append(dtype.itemsize)
        want.extend(dtype.fields[f][1] for f in fields)
    assert values == want
