"""The internal-force kernel with the wide register window (Gauss points with
more support entries than the default window, as RPIM supports of 30
nodes): bit-identical to the default kernel where both apply, and on a
perforated RPIM plate, where it is the kernel the library selects, equal to
the oracle within the FP64 drop-in bar."""
import numpy as np
import pytest

from mel_compare import compare, device_state, layout_from_dump, run_oracle
from melody2d_amd import native
from melody2d_amd.preprocessor import configs
from melody2d_amd.preprocessor.binary import write_binary

pytestmark = pytest.mark.gpu

FIELDS = ["x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param", "x_internal_force", "y_internal_force",
          "x_contact_force", "y_contact_force", "x_current", "y_current"]


def _fields(ctx, nodes):
    return {"b%d.%s" % (b, f): ctx.node_field(b, f, n) for b, n in enumerate(nodes) for f in FIELDS}


@pytest.mark.parametrize("precision", [native.FP64, native.FP32_MIXED])
@pytest.mark.parametrize("name", ["tiny", "tiny_czm", "tiny_platen"])
def test_wide_window_equals_default(name, precision, tmp_path, monkeypatch):
    print("case wide vs default", name, precision, flush=True)
    m = configs.variants()[name]
    path = str(tmp_path / "m.bin")
    write_binary(m, path)
    nodes = [b.n_nodes for b in m.bodies]
    monkeypatch.setenv("MELODY2D_FUSED_STEP", "0")
    ref = native.Context(path, binary=True, precision=precision)
    monkeypatch.setenv("MELODY2D_WIDE_LANES", "1")
    wide = native.Context(path, binary=True, precision=precision)
    prev = 0
    for s in (1, 10, 60):
        for c in (ref, wide):
            c.euler_steps(s - prev)
        prev = s
        a, b = _fields(ref, nodes), _fields(wide, nodes)
        diff = [k for k in a if not np.array_equal(a[k], b[k])]
        assert not diff, "step %d: %s" % (s, diff[:6])
    ref.close()
    wide.close()


def plate(n_nodes=2500):
    m = configs.c3(n_nodes=n_nodes, workers=1)
    b = m.bodies[0]
    # start from the uniaxial 10/s velocity field, so the plate strains
    v = np.stack([10.0 * b.X[:, 0], -0.33 * 10.0 * b.X[:, 1]], axis=1)
    b.velocity = v
    return m


@pytest.mark.parametrize("exact", [False, True])
def test_rpim_plate_matches_oracle(exact, tmp_path, monkeypatch):
    """Balanced sums: within 1e-12 through step 1 and 1e-11 at step 10
    (RPIM node sums cancel more than MLS ones, so rounding-order
    differences grow along the trajectory: ~2e-12 by step 10 on the plate
    with hole borders, 2e-11 by step 50).  The reference's summation order
    (MELODY2D_EXACT_SUMS=1): within 1e-12 through step 50."""
    print("case rpim plate exact=%d" % exact, flush=True)
    if exact:
        monkeypatch.setenv("MELODY2D_EXACT_SUMS", "1")
    m = plate()
    path = str(tmp_path / "plate.bin")
    write_binary(m, path)
    assert max(np.diff(m.bodies[0].gp_ptr)) == 30
    steps = [0, 1, 10, 50] if exact else [0, 1, 10]
    refs, _ = run_oracle(path, steps, str(tmp_path / "oracle"), save_index=-1)
    ctx = native.Context(path, binary=True)
    ctx.set_kernel_timing(["block_force_update", "step_forces"])
    prev = 0
    for s in steps:
        ctx.euler_steps(s - prev)
        prev = s
        nn, nbr, sizes = layout_from_dump(refs[s])
        res = compare(device_state(ctx, nn, nbr, sizes), refs[s])
        tol = 1e-12 if exact or s <= 1 else 1e-11
        bad = {k: v[0] for k, v in res.items() if not v[0] <= tol}
        assert not bad, "step %d: %s" % (s, sorted(bad.items(), key=lambda kv: -kv[1])[:8])
    f = np.abs(ctx.node_field(0, "x_internal_force", nn[0])).max()
    assert f > 0
    kt = ctx.kernel_timing()
    # the launch sequence (wide window), not the two-launch step
    assert kt["block_force_update"][1] == steps[-1] and kt["step_forces"][1] == 0
    ctx.close()


def test_plate_fp32_mixed_force_error_within_bar(tmp_path):
    """The FP32-mixed internal forces of the perforated RPIM plate in the
    uniaxial 10/s field (strain ~1.6e-5, forces small differences of large
    stress terms) are within 1e-5 of FP64's from the same state; the
    displacement increments of 100 further steps too."""
    print("case plate fp32-mixed error", flush=True)
    m = configs.c3(n_nodes=10000, side=0.1 * np.sqrt(0.01), workers=4)
    b = m.bodies[0]
    path = str(tmp_path / "plate.bin")
    write_binary(m, path)
    eps = 10.0 * 1000 * m.dt
    x, y = b.X[:, 0], b.X[:, 1]
    ctx = {p: native.Context(path, binary=True, precision=p) for p in (native.FP64, native.FP32_MIXED)}
    for c in ctx.values():
        for f, v in (("x_disp_param", eps * x), ("y_disp_param", -0.33 * eps * y),
                     ("x_vel_param", 10.0 * x), ("y_vel_param", -0.33 * 10.0 * y)):
            c.set_node_field(0, f, np.ascontiguousarray(v))
        c.eval_forces(native.EVAL_INTERNAL)
    n = b.n_nodes
    f64 = np.concatenate([ctx[native.FP64].node_field(0, f, n) for f in ("x_internal_force", "y_internal_force")])
    f32 = np.concatenate([ctx[native.FP32_MIXED].node_field(0, f, n) for f in ("x_internal_force", "y_internal_force")])
    err = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    print("  internal-force error %.3e" % err, flush=True)
    assert err <= 1e-5, err
    u0 = ctx[native.FP64].node_field(0, "x_disp_param", n)
    for c in ctx.values():
        c.euler_steps(100)
    d64 = ctx[native.FP64].node_field(0, "x_disp_param", n) - u0
    d32 = ctx[native.FP32_MIXED].node_field(0, "x_disp_param", n) - u0
    assert np.linalg.norm(d32 - d64) <= 1e-5 * np.linalg.norm(d64)
    for c in ctx.values():
        c.close()
