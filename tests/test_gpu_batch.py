"""Batched samples in one context (`configs.c4`): independent periodic shear
cells stacked in one model run as one batch of launches per step; every
sample's state is bit for bit the state it reaches alone (the same sample
built alone at the same place, with the batch's time step), in both
precisions, across contact updates."""
import numpy as np
import pytest

from melody2d_amd import native
from melody2d_amd.preprocessor.binary import write_binary
from melody2d_amd.preprocessor.configs import c4_batch

pytestmark = pytest.mark.gpu

FIELDS = ["x_current", "y_current", "x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param",
          "x_internal_force", "y_internal_force", "x_contact_force", "y_contact_force", "x_force", "y_force"]
KW = dict(n_samples=4, grains=(6, 14), seed=2004, workers=4)


def _body_state(ctx, b, n, body_shift):
    d = {f: ctx.node_field(b, f, n) for f in FIELDS}
    ids, st = ctx.contacts(b)
    ids = ids.reshape(-1, 16).copy()
    ids[:, 3] -= body_shift            # master body, relative to the sample's first body
    d["contacts_i"] = ids
    d["contacts_d"] = st
    nb = ctx.neighbours(b).reshape(-1, 3).copy()
    nb[:, 0] -= body_shift
    d["neighbours"] = nb
    r, w = ctx.body_state(b)
    d["rigid"], d["works"] = r, w
    return d


@pytest.mark.parametrize("precision", [native.FP64, native.FP32_MIXED])
def test_batch_samples_equal_alone(precision, tmp_path):
    print("case batch vs alone, precision", precision, flush=True)
    batch = c4_batch(**KW)
    path = str(tmp_path / "batch.bin")
    write_binary(batch, path)
    counts = [b.n_nodes for b in batch.bodies]
    per_sample = [g + 2 for g, _, _ in batch.samples]
    first = np.concatenate([[0], np.cumsum(per_sample)])
    ctx = native.Context(path, binary=True, precision=precision)
    ctx.euler_steps(57)
    contacts = 0
    for s in range(len(per_sample)):
        alone = c4_batch(only=s, dt=batch.dt, **KW)
        p = str(tmp_path / ("alone%d.bin" % s))
        write_binary(alone, p)
        a = native.Context(p, binary=True, precision=precision)
        a.euler_steps(57)
        for k in range(per_sample[s]):
            b = first[s] + k
            want = _body_state(a, k, counts[b], 0)
            got = _body_state(ctx, b, counts[b], first[s])
            contacts += len(want["contacts_d"]) // 18
            diff = [f for f in want if not np.array_equal(want[f], got[f])]
            assert not diff, (s, k, diff)
        a.close()
    assert contacts > 0
    ctx.close()
