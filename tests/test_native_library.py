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


def test_damaged_container_reports_parse_error(tmp_path):
    """The binary container reader checks every record length against the
    bytes left in the file: random bytes and a cut-off container are parse
    errors (before any device call), not allocation failures."""
    import pytest
    import numpy as np
    from melody2d_amd.preprocessor import configs
    from melody2d_amd.preprocessor.binary import write_binary
    rng = np.random.default_rng(5)
    bad = tmp_path / "random.bin"
    bad.write_bytes(rng.integers(0, 256, 4096, dtype=np.uint8).tobytes())
    good = tmp_path / "tiny.bin"
    write_binary(configs.variants()["tiny"], str(good))
    cut = tmp_path / "cut.bin"
    data = good.read_bytes()
    cut.write_bytes(data[: len(data) // 2])
    for p in (bad, cut):
        with pytest.raises(native.MelError) as e:
            native.Context(str(p), binary=True)
        assert e.value.status == 3, p
    with pytest.raises(native.MelError) as e:
        native.Context(str(tmp_path / "missing.bin"), binary=True)
    assert e.value.status == 2


def test_server_entry_points_check_arguments():
    """Argument checks of the job server that need no device."""
    lib = native.library()
    info = native.JobInfo()
    assert lib.mel_server_job(None, 0, ctypes.byref(info)) == 1
    assert lib.mel_server_run(None, 1.0, None) == 1
    assert lib.mel_server_release(None, 0) == 1
    assert lib.mel_get_device_bytes(None, None) == 1
    h = ctypes.c_void_p()
    assert lib.mel_server_create(0, 0, 10, ctypes.byref(h)) == 1      # no budget
    assert lib.mel_server_create(0, 1 << 30, 0, ctypes.byref(h)) == 1  # no slice
    lib.mel_server_destroy(None)


def test_load_options_are_checked_before_the_device():
    """A node-block size outside 1..256 is an argument error, reported before
    the model is read or any device call is made."""
    import pytest
    for bn in (0, 257, 100000):
        with pytest.raises(native.MelError) as e:
            native.Context("unused.bin", binary=True, options={"block_nodes": bn})
        assert e.value.status == 1, bn
    o = native.Options()
    native.library().mel_default_options(ctypes.byref(o))
    assert o.as_dict() == {"block_nodes": -1, "fused_step": -1, "graph_replay": -1, "wide_window": -1}
    assert native.library().mel_autotune(None, 0, 0, 10, None, None, None) == 1


def _records(data):
    """(name, kind, count, payload offset) of every record of a container."""
    import struct
    out, pos = [], 0
    while pos < len(data):
        (ln,) = struct.unpack_from("<I", data, pos)
        name = data[pos + 4:pos + 4 + ln].decode()
        kind, n = struct.unpack_from("<BQ", data, pos + 4 + ln)
        off = pos + 4 + ln + 9
        out.append((name, kind, n, off))
        pos = off + n * (8 if kind == 0 else 4)
    return out


def test_patched_index_records_are_parse_errors(tmp_path):
    """Index records that contradict each other (pointer arrays out of range
    or out of order, node, border, material and shape-function indices out
    of range) are parse errors reported before any device call, never a
    crash of the process."""
    import struct
    import pytest
    from melody2d_amd.preprocessor import configs
    from melody2d_amd.preprocessor.binary import write_binary
    good = tmp_path / "tiny.bin"
    write_binary(configs.variants()["tiny_platen"], str(good))
    data = good.read_bytes()
    recs = {r[0]: r for r in _records(data)}
    cases = [("b0.border_meta", 1, 10 ** 8), ("b0.border_meta", 2, -3), ("b0.border_ptr", 1, 10 ** 8),
             ("b0.border_ptr", 1, 0), ("b0.border_nodes", 3, 10 ** 8), ("b0.gidx", 5, 10 ** 8),
             ("b0.gidx", 7, -1), ("b0.gptr", 3, 10 ** 8), ("b0.didx", 0, 10 ** 8), ("b0.dptr", 2, -5),
             ("b0.info", 1, 99), ("b0.info", 4, 10 ** 8), ("b0.cells", 0, 10 ** 8), ("b0.border_hist_ptr", 1, 10 ** 8),
             ("model.law_materials", 0, 50), ("model.law_ptr", 1, 10 ** 8), ("model.n_bodies", 0, 10 ** 8),
             ("b2.border_meta", 1, 7), ("b2.info", 1, -1)]
    for name, k, v in cases:
        _, kind, n, off = recs[name]
        assert kind == 1 and k < n, name
        bad = bytearray(data)
        struct.pack_into("<i", bad, off + 4 * k, v)
        p = tmp_path / "bad.bin"
        p.write_bytes(bytes(bad))
        with pytest.raises(native.MelError) as e:
            native.Context(str(p), binary=True)
        assert e.value.status == 3, (name, k, v, e.value)
    # the unpatched container still reads (the first device call is where it stops without a GPU)
    assert recs["b0.border_meta"][2] >= 5
