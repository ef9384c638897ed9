"""The C-ABI library loads and exports every entry point include/melody2d.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from melody2d_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "melody2d.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mel_[a-z_0-9]+)\s*\(", txt)))


def test_header_declarations_are_exported():
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(native.exported_symbols()) == names


def test_version_string():
    assert native.library().mel_version().decode().startswith("melody2d")


def test_missing_model_reports_io_error(tmp_path):
    # mel_load parses the files before touching any device
    import pytest
    with pytest.raises(native.MelError) as e:
        native.Context(str(tmp_path / "nowhere"))
    assert e.value.status == 2
