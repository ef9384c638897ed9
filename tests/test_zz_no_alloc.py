"""mel_eval_forces_noalloc (no device allocation; contact state saved in
the library's static device scratch) against mel_eval_forces, through the
C-ABI: the same force fields bit for bit on two calls in a row, the second
on a model with fewer contact elements, the context's device footprint
unchanged by the call, and the state after it untouched (a step after the
evaluation matches a step without it, bit for bit)."""
import numpy as np
import pytest

from mel_compare import device_state, layout_from_dump, run_oracle
from melody2d_amd import native

pytestmark = pytest.mark.gpu

FIELDS = ["x_internal_force", "y_internal_force", "x_contact_force", "y_contact_force", "x_dirichlet_force",
          "y_dirichlet_force", "x_neumann_force", "y_neumann_force", "x_body_force", "y_body_force",
          "x_damping_force", "y_damping_force", "x_force", "y_force"]


def _fields(ctx, nn):
    return {(b, f): ctx.node_field(b, f, n) for b, n in enumerate(nn) for f in FIELDS}


def test_noalloc_eval_equals_allocating_eval(golden_dir, tmp_path):
    for name, steps in (("tiny_holed", 5), ("tiny_czm", 7)):   # the second has fewer contact elements
        print("case noalloc", name, flush=True)
        model, _ = golden_dir(name)
        refs, _ = run_oracle(model, [0], str(tmp_path / name))
        nn, nbr, sizes = layout_from_dump(refs[0])
        a = native.Context(model)   # allocating evaluation
        z = native.Context(model)   # no-allocation evaluation
        plain = native.Context(model)
        for c in (a, z, plain):
            c.euler_steps(steps)
        n_el = sum(len(c) for c in (a.contacts(b)[0] for b in range(len(nn)))) // 16
        assert n_el > 0
        before = z.device_bytes()
        a.eval_forces(native.EVAL_ALL)
        z.eval_forces_noalloc(native.EVAL_ALL)
        assert z.device_bytes() == before
        fa, fz = _fields(a, nn), _fields(z, nn)
        diff = [k for k in fa if not np.array_equal(fa[k], fz[k])]
        assert not diff, (name, diff[:6])
        assert any(np.any(v) for k, v in fz.items() if k[1] in ("x_contact_force", "y_contact_force"))
        # the evaluation left the state as it was
        for c in (z, plain):
            c.euler_steps(3)
        sz, sp = device_state(z, nn, nbr, sizes), device_state(plain, nn, nbr, sizes)
        diff = [k for k in sp if not np.array_equal(np.asarray(sz[k]), np.asarray(sp[k]))]
        assert not diff, (name, diff[:6])
        for c in (a, z, plain):
            c.close()
