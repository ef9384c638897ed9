"""Parity of the HIP library (through the C-ABI) with the oracle.

On each golden model the device and the oracle run from the same files
(with the model's scheme: Euler_step, or Solver_step for the adaptive
ones) and are compared record by record after 0, 1, 2, 10, 50 and 200 steps:
integer records (contact ids, neighbour lists) exactly, floating-point
records within 1e-12 relative (normwise), the FP64 drop-in bar.  The oracle
itself reproduces the reference bit for bit on these models
(test_oracle_golden.py)."""
import os

import numpy as np
import pytest

from conftest import TWO_LAUNCH_VARIANTS, VARIANTS
from mel_compare import compare, device_state, layout_from_dump, run_oracle
from melody2d_amd import native

STEPS = [0, 1, 2, 10, 50, 200]
TOL = 1e-12
# the adaptive schemes' error estimates (Node::Compute_error, Code/Node.h:253-272)
# are differences of two nearly equal displacements, the two-half-step one
# and the one-step predictor: rounding differences of the displacements at
# 1e-16 relative reach them amplified by |u| / |u - u_temp| (about 1e4 here)
ERR_TOL = 1e-7
ERR_RECORDS = (".error", "max_error", "mean_error")
# the time step of the error-controlled schemes follows the error estimate
# (dt / (error / target)^0.2, Code/Solver.h:340-350), so those differences
# feed back into the trajectory: past step 50 their records are held to
# 1e-10; Mass_Scaling also switches every node factor between growing and
# decaying on the same estimate against a threshold (Code/Node.h:278-306),
# and its growth term is a pow() that the device rounds on its own: 1e-7
ADAPTIVE_LATE_TOL = {"tiny_adaptive": 1e-10, "tiny_adaptive_ms": 1e-10, "tiny_mass_scaling": 1e-7,
                     "tiny_deact_adaptive": 1e-10}


# a Monitoring row's entries are single node or body values, each held to
# its own magnitude rather than to the norm of a field
MON_TOL = 1e-9


def _tol(key, step=0, name=""):
    if any(key.endswith(r) for r in ERR_RECORDS):
        return ERR_TOL
    if key == "monitor":
        return MON_TOL
    return ADAPTIVE_LATE_TOL[name] if name in ADAPTIVE_LATE_TOL and step > 50 else TOL

pytestmark = pytest.mark.gpu


def _run_pair(model_dir, steps, tmp, binary_path=None):
    refs, _ = run_oracle(model_dir, steps, str(tmp / "oracle"))
    ctx = native.Context(binary_path, binary=True) if binary_path else native.Context(model_dir)
    # models with MONITORING entries: the row of each compared step
    monitor = "monitor" in refs[max(steps)]
    if monitor:
        ctx.set_monitoring(True)
    out = {}
    prev = 0
    for s in steps:
        ctx.solver_steps(s - prev)
        prev = s
        ctx.synchronize()
        nn, nbr, sizes = layout_from_dump(refs[s])
        dev = device_state(ctx, nn, nbr, sizes)
        if monitor and s > 0:
            rows = ctx.monitoring()
            assert len(rows) == s - (steps[steps.index(s) - 1] if steps.index(s) > 0 else 0)
            dev["monitor"] = rows[-1]
        out[s] = (dev, refs[s])
    ctx.close()
    return out


@pytest.mark.parametrize("name", VARIANTS)
def test_device_matches_oracle(name, golden_dir, tmp_path):
    model, _ = golden_dir(name)
    print("case", name, flush=True)
    # the deactivation variant also at the contact update that stops a body
    # and within the period after it
    steps = sorted(STEPS + [80, 90]) if name == "tiny_deact" else STEPS
    for s, (dev, ref) in _run_pair(model, steps, tmp_path).items():
        res = compare(dev, ref)
        assert len(res) > 20
        bad = {k: v[0] for k, v in res.items() if not v[0] <= _tol(k, s, name)}
        if "monitor" in bad:
            d, r = np.asarray(dev["monitor"]), np.asarray(ref["monitor"])
            e = [abs(d[i] - r[i]) / max(abs(r[i]), 1e-300) for i in range(len(r))]
            print("monitor entries", [(int(i), d[i], r[i]) for i in np.argsort(e)[::-1][:4]], flush=True)
        assert not bad, "step %d: %s" % (s, sorted(bad.items(), key=lambda kv: -kv[1])[:8])


def test_internal_forces_bitwise_on_contact_free_start(golden_dir, tmp_path, monkeypatch):
    """With MELODY2D_EXACT_SUMS=1 internal forces are gathered per region in
    Gauss-point order as the reference does: at step 1 (same displacement
    parameters on both sides) they are bit-identical wherever
    pow(J, -5/3) rounds as glibc does."""
    monkeypatch.setenv("MELODY2D_EXACT_SUMS", "1")
    model, _ = golden_dir("tiny")
    dev, ref = _run_pair(model, [1], tmp_path)[1]
    for b in range(2):
        for f in ("x_internal_force", "y_internal_force"):
            k = "b%d.%s" % (b, f)
            eq = np.mean(dev[k] == ref[k])
            assert eq > 0.9, (k, eq)


@pytest.mark.parametrize("name", ["tiny_bc", "tiny_platen"])
def test_split_node_blocks_match_oracle(name, golden_dir, tmp_path, monkeypatch):
    """Bodies cut into several node blocks (64 nodes each here) read
    displacement parameters owned by sibling blocks: results are the same as
    with one block per body."""
    monkeypatch.setenv("MELODY2D_BLOCK_NODES", "64")
    model, _ = golden_dir(name)
    print("case split", name, flush=True)
    for s, (dev, ref) in _run_pair(model, [1, 10, 50], tmp_path).items():
        res = compare(dev, ref)
        bad = {k: v[0] for k, v in res.items() if not v[0] <= TOL}
        assert not bad, "step %d: %s" % (s, sorted(bad.items(), key=lambda kv: -kv[1])[:8])


@pytest.mark.parametrize("name", ["tiny_bc", "tiny_czm"])
def test_exact_sum_order_matches_oracle(name, golden_dir, tmp_path, monkeypatch):
    """The reference-order gather (contiguous Gauss-point chunks, per-region
    partial sums) meets the same bar as the default balanced order."""
    monkeypatch.setenv("MELODY2D_EXACT_SUMS", "1")
    model, _ = golden_dir(name)
    print("case exact", name, flush=True)
    for s, (dev, ref) in _run_pair(model, [1, 10, 50, 200], tmp_path).items():
        res = compare(dev, ref)
        bad = {k: v[0] for k, v in res.items() if not v[0] <= TOL}
        assert not bad, "step %d: %s" % (s, sorted(bad.items(), key=lambda kv: -kv[1])[:8])


@pytest.mark.parametrize("name", TWO_LAUNCH_VARIANTS)
def test_two_launch_step_equals_launch_sequence(name, golden_dir, tmp_path, monkeypatch):
    """The step in two launches (contact elements beside the internal forces,
    then rigid bodies, node-block finish and BC update) and the launch
    sequence (MELODY2D_FUSED_STEP=0) run the same device routines in the same
    orders: every record is bit-identical after 1, 10 and 200 steps."""
    _two_launch_vs_sequence(name, golden_dir, tmp_path, monkeypatch, native.FP64)


@pytest.mark.parametrize("name", ["tiny_platen", "tiny_czm", "tiny_bc"])
def test_two_launch_step_equals_launch_sequence_fp32_mixed(name, golden_dir, tmp_path, monkeypatch):
    """The same in the FP32-mixed form, whose internal-force pass sums one
    chunk while the next chunk's Gauss points write the other contribution
    buffer, in both paths."""
    _two_launch_vs_sequence(name, golden_dir, tmp_path, monkeypatch, native.FP32_MIXED)


def _two_launch_vs_sequence(name, golden_dir, tmp_path, monkeypatch, precision):
    model, _ = golden_dir(name)
    print("case two-launch vs sequence", name, precision, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, nbr, sizes = layout_from_dump(refs[0])
    two = native.Context(model, precision=precision)
    two.set_kernel_timing(["step_forces"])
    monkeypatch.setenv("MELODY2D_FUSED_STEP", "0")
    seq = native.Context(model, precision=precision)
    seq.set_kernel_timing(["step_forces"])
    prev = 0
    for s in (1, 10, 200):
        two.euler_steps(s - prev)
        seq.euler_steps(s - prev)
        prev = s
        a, b = device_state(two, nn, nbr, sizes), device_state(seq, nn, nbr, sizes)
        assert a.keys() == b.keys()
        diff = [k for k in a if not np.array_equal(np.asarray(a[k]), np.asarray(b[k]))]
        assert not diff, "step %d: %s" % (s, diff[:8])
    assert two.kernel_timing()["step_forces"][1] == 200
    assert seq.kernel_timing()["step_forces"][1] == 0
    two.close()
    seq.close()


@pytest.mark.parametrize("name", ["tiny_platen", "tiny_czm", "tiny_damped"])
def test_graph_replays_equal_launches(name, golden_dir, tmp_path, monkeypatch):
    """Runs of steps between contact updates replayed from captured graphs
    give the same state as the same launches issued one by one
    (MELODY2D_GRAPHS=0), replay after replay, with API reads and single steps
    (eager launches) between the replays."""
    model, _ = golden_dir(name)
    print("case graphs", name, flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, nbr, sizes = layout_from_dump(refs[0])
    gr = native.Context(model)
    monkeypatch.setenv("MELODY2D_GRAPHS", "0")
    ea = native.Context(model)
    for chunk in (10, 1, 25, 10, 3, 40, 10):
        for c in (gr, ea):
            c.euler_steps(chunk)
        a, b = device_state(gr, nn, nbr, sizes), device_state(ea, nn, nbr, sizes)
        diff = [k for k in a if not np.array_equal(np.asarray(a[k]), np.asarray(b[k]))]
        assert not diff, "after a run of %d: %s" % (chunk, diff[:8])
    gr.close()
    ea.close()
