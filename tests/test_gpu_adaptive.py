"""The adaptive schemes of Solver_step (Code/Solver.h:98-486) on the device,
beyond the golden comparison of test_gpu_parity.py: the in-memory checkpoint
carries the time step and the mass-scaling factors, rejected attempts leave
no trace, and the FP32-mixed form runs the schemes."""
import numpy as np
import pytest

from conftest import ADAPTIVE_VARIANTS
from mel_compare import device_state, layout_from_dump, run_oracle
from melody2d_amd import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ADAPTIVE_VARIANTS)
def test_state_roundtrip_continues_bitwise(name, golden_dir, tmp_path):
    """A context restored from another's saved state (mid-run, between two
    accepted steps) continues bit for bit like the one it was saved from."""
    model, _ = golden_dir(name)
    print("case adaptive roundtrip", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, nbr, sizes = layout_from_dump(refs[0])
    a = native.Context(model)
    a.solver_steps(37)
    blob = a.save_state()
    b = native.Context(model)
    b.load_state(blob)
    assert b.solver_info().dt == a.solver_info().dt
    a.solver_steps(41)
    b.solver_steps(41)
    da, db = device_state(a, nn, nbr, sizes), device_state(b, nn, nbr, sizes)
    diff = [k for k in da if not np.array_equal(np.asarray(da[k]), np.asarray(db[k]))]
    assert not diff, diff[:8]
    # the error-controlled schemes rejected some attempts on the way (the
    # mass-ratio one only rejects later, once the scaled masses have grown)
    assert a.solver_info().attempts >= 78
    if name in ("tiny_adaptive", "tiny_mass_scaling"):
        assert a.solver_info().attempts > 78
    a.close()
    b.close()


@pytest.mark.parametrize("name", ADAPTIVE_VARIANTS)
def test_fp32_mixed_runs_the_scheme(name, golden_dir, tmp_path):
    """The FP32-mixed form runs the adaptive scheme: 100 accepted steps with
    finite states, the same number of attempts' worth of simulated time as
    the FP64 form to 1e-3, and positions within 1e-5 of it."""
    model, _ = golden_dir(name)
    print("case adaptive fp32", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, _, _ = layout_from_dump(refs[0])
    c64 = native.Context(model)
    c32 = native.Context(model, precision=native.FP32_MIXED)
    c64.solver_steps(100)
    c32.solver_steps(100)
    t64, t32 = c64.get_info().time, c32.get_info().time
    assert abs(t32 - t64) <= 1e-3 * t64, (t32, t64)
    for b, n in enumerate(nn):
        x64 = c64.node_field(b, "x_current", n)
        x32 = c32.node_field(b, "x_current", n)
        assert np.all(np.isfinite(x32))
        assert np.linalg.norm(x32 - x64) <= 1e-5 * np.linalg.norm(x64), (b, np.linalg.norm(x32 - x64))
    c64.close()
    c32.close()
