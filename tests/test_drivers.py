"""The reference's driver over the C-ABI (melody2d_amd/examples/main_hip.cpp,
shown in INTEGRATION.md) and over the C++ façade (main_facade.cpp with
include/melody2d.hpp): they compile against the headers alone, and on a GPU
they write the same DYNAMIC files as the Python binding after the same
steps."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "melody2d_amd", "examples")


def _doc_block(name):
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```cpp\n(// %s\.cpp .*?)```" % re.escape(name), txt, flags=re.S)
    assert m, name
    return m.group(1)


@pytest.mark.parametrize("name", ["main_hip", "main_facade"])
def test_documented_driver_is_the_example_and_compiles(name, tmp_path):
    src = open(os.path.join(EX, name + ".cpp")).read()
    assert _doc_block(name) == src
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(EX, name + ".cpp"), "-L", os.path.join(ROOT, "melody2d_amd"), "-lmelody2d_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "melody2d_amd"), "-o", exe], check=True)
    # without model files the driver stops at the load with the library's message
    r = subprocess.run([exe, "0", "1", "1"], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 1
    assert "STATIC_CONTROL.asc" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["main_hip", "main_facade"])
def test_driver_writes_what_the_binding_writes(name, golden_dir, tmp_path):
    from melody2d_amd import native
    print("case driver", name, flush=True)
    exe = os.path.join(EX, name)
    assert os.path.exists(exe), "build the examples first (melody2d_amd.build.build_all)"
    model, _ = golden_dir("tiny_platen")
    run = tmp_path / "run"
    shutil.copytree(model, str(run))
    r = subprocess.run([exe, "0", "2", "25"], cwd=str(run), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with native.Context(model) as ctx:
        ctx.euler_steps(25)
        ctx.synchronize()
        ctx.write_dynamic(str(tmp_path / "one.asc"))
        ctx.euler_steps(25)
        ctx.synchronize()
        ctx.write_dynamic(str(tmp_path / "two.asc"))
    assert open(str(run / "DYNAMIC_00001.asc")).read() == open(str(tmp_path / "one.asc")).read()
    assert open(str(run / "DYNAMIC_00002.asc")).read() == open(str(tmp_path / "two.asc")).read()


@pytest.mark.gpu
def test_python_facade_matches_the_binding(golden_dir):
    """melody2d_amd.solver (Model / Body / Solver) drives the same context
    calls: the same state after the same steps."""
    import numpy as np
    from melody2d_amd import native
    from melody2d_amd.solver import Model, Solver
    print("case python facade", flush=True)
    model_dir, _ = golden_dir("tiny_czm")
    with Model.load(model_dir) as m, native.Context(model_dir) as ctx:
        s = Solver()
        for _ in range(30):
            s.euler_step(m)
        ctx.euler_steps(30)
        assert m.number_iteration == 30 and m.time == ctx.get_info().time
        for b in m.bodies:
            for f in ("x_disp", "y_vel_param", "x_contact_force"):
                assert np.array_equal(b.field(f), ctx.node_field(b.index, f, b.nb_nodes)), (b.index, f)
            ids, st = b.contacts()
            ids2, st2 = ctx.contacts(b.index)
            assert np.array_equal(ids, ids2) and np.array_equal(st, st2)
        t, w, e = m.scalars()
        t2, w2, e2 = ctx.scalars()
        assert t == t2 and np.array_equal(w, w2) and np.array_equal(e, e2)
