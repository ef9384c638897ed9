"""FP32-mixed form against the FP64 form, from the same state.

Both forms are loaded from the same golden model; the FP64 context runs 200
steps, its displacement and velocity parameters are copied into the FP32
context, and one step of each evaluates internal forces from identical
displacements (contact and rigid-body state are not copied, so only the
internal forces are compared there).  Bar: relative force error
||f32 - f64|| / ||f64|| <= 1e-5, for each body.  A second test steps both
forms from the same initial state and compares positions."""
import numpy as np
import pytest

from conftest import VARIANTS
from mel_compare import layout_from_dump, run_oracle
from melody2d_amd import native

pytestmark = pytest.mark.gpu

TOL = 1e-5
COPIED = ["x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param"]


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / n if n > 0 else np.linalg.norm(a)


@pytest.mark.parametrize("name", ["tiny", "tiny_bc", "tiny_platen", "tiny_czm"])
def test_fp32_internal_forces_within_1e5(name, golden_dir, tmp_path):
    model, _ = golden_dir(name)
    print("case fp32", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, _, _ = layout_from_dump(refs[0])
    c64 = native.Context(model)
    c32 = native.Context(model, precision=native.FP32_MIXED)
    assert c32.get_info().precision == native.FP32_MIXED
    c64.euler_steps(200)
    c64.synchronize()
    for b, n in enumerate(nn):
        for f in COPIED:
            c32.set_node_field(b, f, c64.node_field(b, f, n))
    c64.euler_steps(1)
    c32.euler_steps(1)
    checked = 0
    for b, n in enumerate(nn):
        f64 = np.concatenate([c64.node_field(b, "x_internal_force", n), c64.node_field(b, "y_internal_force", n)])
        f32 = np.concatenate([c32.node_field(b, "x_internal_force", n), c32.node_field(b, "y_internal_force", n)])
        if np.linalg.norm(f64) == 0:
            continue
        err = _rel(f32, f64)
        print("  body %d internal-force rel err %.3e" % (b, err), flush=True)
        assert err <= TOL, (b, err)
        checked += 1
    assert checked > 0
    c64.close()
    c32.close()


def test_fp32_runs_every_golden_model(golden_dir, tmp_path):
    """The FP32-mixed form loads and steps every golden model; after 50
    steps its displacements agree with the FP64 form's to 1e-5 relative."""
    for name in VARIANTS:
        model, _ = golden_dir(name)
        refs, _ = run_oracle(model, [0], str(tmp_path / ("o_" + name)))
        nn, _, _ = layout_from_dump(refs[0])
        c64 = native.Context(model)
        c32 = native.Context(model, precision=native.FP32_MIXED)
        c64.euler_steps(50)
        c32.euler_steps(50)
        for b, n in enumerate(nn):
            x64 = c64.node_field(b, "x_current", n)
            x32 = c32.node_field(b, "x_current", n)
            assert np.all(np.isfinite(x32)), (name, b)
            assert _rel(x32, x64) <= TOL, (name, b, _rel(x32, x64))
        c64.close()
        c32.close()


@pytest.mark.parametrize("name", ["tiny_platen", "tiny_frictionless", "tiny_cohesion", "tiny_damped",
                                  "tiny_twoslopes", "tiny_czm"])
def test_fp32_contact_law_forces_within_1e5(name, golden_dir, tmp_path):
    """The FP32-mixed form applies the contact laws in FP32 (gaps and
    projections stay FP64): from one FP64 state, its contact forces are within
    1e-5 of the FP64 form's, and not identical to them."""
    model, _ = golden_dir(name)
    print("case fp32 laws", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, _, _ = layout_from_dump(refs[0])
    c64 = native.Context(model)
    c32 = native.Context(model, precision=native.FP32_MIXED)

    def forces(c):
        return np.concatenate([np.concatenate([c.node_field(b, "x_contact_force", n),
                                               c.node_field(b, "y_contact_force", n)]) for b, n in enumerate(nn)])

    # the first of these states at which contacts carry a force
    done = 0
    for k in (0, 1, 2, 5, 10, 20, 50, 100, 200):
        c64.euler_steps(k - done)
        done = k
        c64.eval_forces(native.EVAL_CONTACT)
        f64 = forces(c64)
        if np.linalg.norm(f64) > 0:
            break
    assert np.linalg.norm(f64) > 0
    c32.load_state(c64.save_state())
    c32.eval_forces(native.EVAL_CONTACT)
    f32 = forces(c32)
    print("  state after %d steps" % done, flush=True)
    err = _rel(f32, f64)
    print("  contact-force rel err %.3e" % err, flush=True)
    assert 0 < err <= TOL
    c64.close()
    c32.close()
