"""The C-ABI library loads on the CPU and exports every declared entry point."""

import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(art_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_python_binding():
    from art import _hip

    assert declared_functions() == set(_hip.exported_symbols())


def test_library_exports_every_declared_symbol():
    from art import _hip

    lib = _hip.load()  # no GPU call: loading and symbol lookup only
    assert lib.art_version() == 1
    out = subprocess.run(["nm", "-D", "--defined-only", str(_hip.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b(art_[a-z0-9_]+)\b", out))
    missing = declared_functions() - exported
    assert not missing, f"not exported: {sorted(missing)}"


@pytest.mark.parametrize("d,degree,bias,last,p", [
    (2, 2, 1, 0, 6), (8, 2, 1, 0, 45), (12, 2, 1, 0, 91), (16, 1, 1, 1, 17), (3, 1, 0, 0, 3),
])
def test_feature_count_is_host_only(d, degree, bias, last, p):
    from art import _hip

    assert _hip.load().art_feature_count(d, degree, bias, last) == p
    assert _hip.load().art_feature_count(3, 2, 1, 1) == -1


@pytest.mark.parametrize("p,parts,doubles", [
    # Two slots of parts x stride doubles; the stride holds the longest
    # exchanged vector -- both sides' packed (p+1)(p+2)/2 systems and a flag,
    # 3 + 31 statistics, 8 trial losses -- rounded up to 32.
    (6, 8, 2 * 8 * 64), (1, 2, 2 * 2 * 64), (16, 4, 2 * 4 * 320), (32, 2, 2 * 2 * 1152),
])
def test_split_exchange_space_is_host_only(p, parts, doubles):
    from art import _hip

    lib = _hip.load()
    assert lib.art_split_part_doubles(p, parts) == doubles
    assert lib.art_split_part_doubles(0, 2) == -1 and lib.art_split_part_doubles(33, 2) == -1
    assert lib.art_split_part_doubles(6, 0) == -1


def test_split_parts_entry_refuses_bad_arguments_before_any_device_call():
    import ctypes

    from art import _hip

    lib = _hip.load()
    a = _hip.SplitArgs()
    null = ctypes.c_void_p()
    assert lib.art_split_max_parts(None, 4) == 1
    assert lib.art_split_optimize_parts(None, 4, 2, null, null, null, null) == _hip.ERR_ARG
    # parts outside [1, 8] or missing exchange buffers
    assert lib.art_split_optimize_parts(ctypes.byref(a), 4, 9, null, null, null,
                                        null) == _hip.ERR_ARG
    assert lib.art_split_optimize_parts(ctypes.byref(a), 4, 2, null, null, null,
                                        null) == _hip.ERR_ARG
