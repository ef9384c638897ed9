"""Monitoring (Code/Monitoring.h:8-559) on the device beyond the golden
comparison of test_gpu_parity.py (which compares tiny_monitor's rows with
the oracle's): the two-launch step forms the rows too, Write_monitoring's
text layout, NO_MONITORING."""
import os

import numpy as np
import pytest

from melody2d_amd import native

pytestmark = pytest.mark.gpu


def test_rows_of_the_two_launch_step(golden_dir):
    """A model without MONITORING entries: every step gives the six leading
    scalars (iteration, time, time step, and zeros for the adaptive
    schemes' neval, max_error, mean_error under Euler_step)."""
    model, _ = golden_dir("tiny_platen")
    print("case monitoring two-launch", flush=True)
    ctx = native.Context(model)      # a model the two-launch step runs
    ctx.set_monitoring(True)
    ctx.euler_steps(25)
    rows = ctx.monitoring()
    assert len(rows) == 25
    dt = rows[0][2]
    t = 0.0
    for k, r in enumerate(rows):
        t = t + dt
        assert len(r) == 6
        assert r[0] == k + 1 and r[2] == dt and not np.any(r[3:])
        assert r[1] == t
    assert rows[-1][1] == ctx.get_info().time
    assert ctx.monitoring() == []
    ctx.close()


def test_write_monitoring_layout(golden_dir, tmp_path):
    """Write_monitoring's text: ten significant digits, each value followed
    by a space, one row per line, appended."""
    model, _ = golden_dir("tiny_monitor")
    print("case monitoring text", flush=True)
    a, b = native.Context(model), native.Context(model)
    for c in (a, b):
        c.set_monitoring(True)
        c.euler_steps(12)
    path = str(tmp_path / "MONITORING.asc")
    a.write_monitoring(path)
    rows = b.monitoring()
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[-1] == "" and len(lines) == 13
    for line, r in zip(lines, rows):
        assert line == "".join("%.10g " % v for v in r)
    a.euler_steps(3)
    a.write_monitoring(path)
    with open(path) as f:
        assert len(f.read().split("\n")) == 16
    a.close()
    b.close()


def test_no_monitoring_refuses(golden_dir):
    model, _ = golden_dir("tiny_nomonitor")
    print("case monitoring NO_MONITORING", flush=True)
    ctx = native.Context(model)
    with pytest.raises(native.MelError):
        ctx.set_monitoring(True)
    ctx.close()
