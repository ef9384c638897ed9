"""Checkpoints through the C-ABI.

* In memory (mel_save_state / mel_load_state): a context restored from a
  state saved after k steps continues exactly as the original does, bit for
  bit, whether it is a new context or the same one rewound; a state that
  does not fit the model is refused before anything is written.
* Across precisions: the FP64 state loaded into an FP32-mixed context gives
  internal forces within 1e-5 of FP64's from the same state.
* On disk (mel_write_dynamic, the DYNAMIC_nnnnn.asc layout of
  Code/IO.h:2351-2636): the library and the oracle restart from the file the
  library wrote (Load_dynamic, Code/IO.h:1871-2343, then the initialisation
  block of Code/main.cpp:103-126) and agree within 1e-12 through 50 more
  steps; the restarted kinematics are the written ones, bit for bit."""
import os
import shutil

import numpy as np
import pytest

from conftest import RESTARTS, RESTART_STEP0_SKIP, extract_restart
from mel_compare import compare, device_state, layout_from_dump, run_oracle
from melody2d_amd import native

pytestmark = pytest.mark.gpu


def _state(ctx, layout):
    return device_state(ctx, *layout)


def _diff(a, b):
    return [k for k in a if not np.array_equal(np.asarray(a[k]), np.asarray(b[k]))]


@pytest.mark.parametrize("name", ["tiny_platen", "tiny_damped", "tiny_periodic"])
def test_state_roundtrip_continues_bitwise(name, golden_dir, tmp_path):
    model, _ = golden_dir(name)
    print("case state", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "o"))
    lay = layout_from_dump(refs[0])
    a = native.Context(model)
    a.euler_steps(23)
    blob = a.save_state()
    a.euler_steps(37)
    want = _state(a, lay)
    b = native.Context(model)
    b.load_state(blob)
    b.euler_steps(37)
    assert not _diff(want, _state(b, lay))
    # the same context, rewound
    a.load_state(blob)
    a.euler_steps(37)
    assert not _diff(want, _state(a, lay))
    # a state of another model is refused, and the context is unchanged
    other, _ = golden_dir("tiny_holed")
    c = native.Context(other)
    with pytest.raises(native.MelError):
        c.load_state(blob)
    with pytest.raises(native.MelError):
        a.load_state(blob[:100])
    assert not _diff(want, _state(a, lay))
    for x in (a, b, c):
        x.close()


@pytest.mark.parametrize("name", ["tiny_bc", "tiny_holed"])
def test_fp64_state_in_fp32_mixed_context(name, golden_dir, tmp_path):
    model, _ = golden_dir(name)
    print("case cross-precision state", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "o"))
    counts = layout_from_dump(refs[0])[0]
    a = native.Context(model)
    a.euler_steps(50)
    blob = a.save_state()
    b = native.Context(model, precision=native.FP32_MIXED)
    b.load_state(blob)
    a.eval_forces(native.EVAL_INTERNAL)
    b.eval_forces(native.EVAL_INTERNAL)
    num = den = 0.0
    for body, nn in enumerate(counts):
        for f in ("x_internal_force", "y_internal_force"):
            x, y = a.node_field(body, f, nn), b.node_field(body, f, nn)
            num += float((x - y) @ (x - y))
            den += float(x @ x)
    assert den > 0 and (num / den) ** 0.5 <= 1e-5, (num / den) ** 0.5
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["tiny_czm", "tiny_bc", "tiny_holed"])
def test_write_dynamic_restart_matches_oracle(name, golden_dir, tmp_path):
    model, _ = golden_dir(name)
    print("case restart", name, flush=True)
    d = str(tmp_path / "model")
    shutil.copytree(model, d)
    a = native.Context(d)
    a.euler_steps(30)
    refs0, _ = run_oracle(d, [0], str(tmp_path / "o0"))
    lay = layout_from_dump(refs0[0])
    written = _state(a, lay)
    a.write_dynamic(os.path.join(d, "DYNAMIC_00001.asc"))
    a.close()
    refs, _ = run_oracle(d, [0, 1, 50], str(tmp_path / "o1"), save_index=1)
    b = native.Context(d, save_index=1)
    got = _state(b, lay)
    for f in ("x_current", "y_current", "x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param"):
        for body in range(len(lay[0])):
            k = "b%d.%s" % (body, f)
            assert np.array_equal(got[k], written[k]), k
    prev = 0
    for s in (0, 1, 50):
        b.euler_steps(s - prev)
        prev = s
        res = compare(_state(b, lay), refs[s], skip=RESTART_STEP0_SKIP if s == 0 else ())
        bad = sorted(((k, v[0]) for k, v in res.items() if not v[0] <= 1e-12), key=lambda kv: -kv[1])
        assert not bad, (s, bad[:6])
    b.close()


@pytest.mark.parametrize("name", RESTARTS)
def test_restart_golden_matches_oracle(name, tmp_path):
    """Restart from the golden restart file (written by the oracle with the
    library's writer, restarted by the reference for the pin): the device
    and the oracle agree within 1e-12."""
    model, _ = extract_restart(name, str(tmp_path / "g"))
    print("case restart golden", name, flush=True)
    steps = [0, 1, 10, 50]
    refs, _ = run_oracle(model, steps, str(tmp_path / "o"), save_index=1)
    lay = layout_from_dump(refs[0])
    b = native.Context(model, save_index=1)
    prev = 0
    for s in steps:
        b.euler_steps(s - prev)
        prev = s
        res = compare(_state(b, lay), refs[s], skip=RESTART_STEP0_SKIP if s == 0 else ())
        bad = sorted(((k, v[0]) for k, v in res.items() if not v[0] <= 1e-12), key=lambda kv: -kv[1])
        assert not bad, (s, bad[:6])
    b.close()


@pytest.mark.parametrize("name", __import__("conftest").STRESSES)
def test_nodal_stresses_match_reference(name, tmp_path):
    """mel_compute_nodal_stresses (Body::Compute_nodal_stresses,
    Code/Body.h:2310-2519) against the reference's stress fields at steps 0,
    10 and 50 of the same run: within 1e-11 of each field's largest value
    (the trajectory agrees to ~1e-12 and pow rounds on the device as on the
    host to within an ulp)."""
    from conftest import extract_golden, golden_steps
    from mel_dump import read_dump
    print("case nodal stresses", name, flush=True)
    model, ref = extract_golden(name + "_stress", str(tmp_path / "g"))
    ctx = native.Context(model)
    done = 0
    for s in golden_steps(ref):
        ctx.euler_steps(s - done)
        done = s
        R = read_dump(os.path.join(ref, "step_%06d.bin" % s))
        got = ctx.nodal_stresses()
        for b, a in enumerate(got):
            want = R["b%d.stress" % b].reshape(19, -1)
            assert a.shape == want.shape
            for f in range(19):
                scale = max(np.abs(want[f]).max(), 1e-300)
                err = np.abs(a[f] - want[f]).max() / scale
                assert err <= 1e-11, (s, b, native.STRESS_FIELDS[f], err)
    ctx.close()


def _read_vtk(path):
    lines = open(path).read().split("\n")
    out = {"fields": {}}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("POINTS"):
            n = int(ln.split()[1])
            out["points"] = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            i += n
        elif ln.startswith("POLYGONS"):
            n = int(ln.split()[1])
            out["cells"] = np.array([[int(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            i += n
        elif ln.startswith("SCALARS"):
            name, comps = ln.split()[1], int(ln.split()[3])
            npts = len(out["points"])
            out["fields"][name] = np.array([[float(v) for v in lines[i + 2 + k].split()] for k in range(npts)])
            assert out["fields"][name].shape == (npts, comps)
            i += 1 + npts
        i += 1
    return out


@pytest.mark.parametrize("name", ["tiny_platen", "tiny_periodic"])
def test_graphic_file_holds_the_state(name, golden_dir, tmp_path):
    """mel_write_graphic (Write_graphic, Code/Graphic.h:21-946): a VTK
    POLYDATA file whose points are every active node (periodic positions
    wrapped into the period, seam cells drawn on both sides), whose triangles
    reference existing points, and whose fields are the state's (6
    significant digits, as the reference prints them)."""
    print("case graphic", name, flush=True)
    model, _ = golden_dir(name)
    ctx = native.Context(model)
    ctx.euler_steps(30)
    path = str(tmp_path / "GRAPHIC_00001.vtk")
    ctx.write_graphic(path, number_print=1)
    G = _read_vtk(path)
    info = ctx.get_info()
    sizes = [ctx.body_sizes(b)[0] for b in range(info.n_bodies)]
    assert len(G["points"]) >= sum(sizes)
    assert G["cells"][:, 0].max() == 3 and G["cells"][:, 1:].max() < len(G["points"])
    stress = ctx.nodal_stresses()
    F = G["fields"]
    assert "03_Displacement" in F and "20_Von_Mises_stress" in F and "32_Internal_work" not in F
    for b, n in enumerate(sizes):
        # a body's nodes, then the copies of its seam cells' vertices
        first = int(np.argmax(F["00_Body_index"][:, 0] == b))
        body = F["00_Body_index"][first:first + n, 0]
        assert np.all(body == b)
        for name_f, want in (("03_Displacement", np.stack([ctx.node_field(b, "x_disp", n), ctx.node_field(b, "y_disp", n)], 1)),
                             ("08_Contact_force", np.stack([ctx.node_field(b, "x_contact_force", n),
                                                            ctx.node_field(b, "y_contact_force", n)], 1)),
                             ("20_Von_Mises_stress", stress[b][13][:, None]),
                             ("21_Major_principal_stress", stress[b][0][:, None])):
            got = F[name_f][first:first + n]
            w = np.where(np.abs(want) < 1e-16, 0., want)
            assert np.allclose(got, w, rtol=1e-5, atol=1e-300), (b, name_f)
    ctx.close()
