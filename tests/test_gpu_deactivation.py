"""Deactivation rules (Update_status, Code/Proximity.h:550-1007) on the
device, beyond the golden comparison of test_gpu_parity.py (which follows
tiny_deact through the contact update that stops its upper disc): the
in-memory checkpoint carries statuses, damage and the stopped body's
contact sends; the DYNAMIC writer records the status, and a run restarted
from that file goes on like the oracle restarted from the same file."""
import os
import shutil

import numpy as np
import pytest

from conftest import RESTART_STEP0_SKIP
from mel_compare import compare, device_state, layout_from_dump, run_oracle
from melody2d_amd import native

pytestmark = pytest.mark.gpu


def test_state_roundtrip_across_deactivation(golden_dir, tmp_path):
    """A context restored from a state saved before the upper disc stops
    continues bit for bit like the one it was saved from, through the
    deactivation at step 80 and the period after it."""
    model, _ = golden_dir("tiny_deact")
    print("case deactivation roundtrip", flush=True)
    refs, _ = run_oracle(model, [0], str(tmp_path / "oracle"))
    nn, nbr, sizes = layout_from_dump(refs[0])
    a = native.Context(model)
    a.euler_steps(75)
    assert a.body_status(1)[0] == 0
    blob = a.save_state()
    b = native.Context(model)
    b.load_state(blob)
    a.euler_steps(45)
    b.euler_steps(45)
    assert a.body_status(1)[0] == 1 and b.body_status(1)[0] == 1
    da, db = device_state(a, nn, nbr, sizes), device_state(b, nn, nbr, sizes)
    diff = [k for k in da if not np.array_equal(np.asarray(da[k]), np.asarray(db[k]))]
    assert not diff, diff[:8]
    # a state saved after the deactivation restores it too
    blob = a.save_state()
    c = native.Context(model)
    c.load_state(blob)
    a.euler_steps(20)
    c.euler_steps(20)
    da, dc = device_state(a, nn, nbr, sizes), device_state(c, nn, nbr, sizes)
    diff = [k for k in da if not np.array_equal(np.asarray(da[k]), np.asarray(dc[k]))]
    assert not diff, diff[:8]
    for x in (a, b, c):
        x.close()


def test_restart_from_written_status(golden_dir, tmp_path):
    """Write_dynamic after the deactivation writes the body as inactive; the
    device restarted from that file and the oracle restarted from it agree
    over 50 further steps (1e-12)."""
    src, _ = golden_dir("tiny_deact")
    model = str(tmp_path / "model")
    shutil.copytree(src, model)
    print("case deactivation restart", flush=True)
    ctx = native.Context(model)
    ctx.euler_steps(100)
    ctx.synchronize()
    dyn = os.path.join(model, "DYNAMIC_00007.asc")
    ctx.write_dynamic(dyn)
    ctx.close()
    with open(dyn) as f:
        assert "inactive" in f.read()
    refs, _ = run_oracle(model, [0, 1, 50], str(tmp_path / "oracle"), save_index=7)
    d = native.Context(model, save_index=7)
    prev = 0
    for s in (0, 1, 50):
        d.euler_steps(s - prev)
        prev = s
        d.synchronize()
        nn, nbr, sizes = layout_from_dump(refs[s])
        # (a restart's step 0 carries no forces, pressures or energies,
        # as for every restart set)
        res = compare(device_state(d, nn, nbr, sizes), refs[s], skip=RESTART_STEP0_SKIP if s == 0 else ())
        bad = {k: v[0] for k, v in res.items() if not v[0] <= 1e-12}
        assert not bad, "step %d: %s" % (s, sorted(bad.items(), key=lambda kv: -kv[1])[:8])
    assert d.body_status(1)[0] == 1
    d.close()
