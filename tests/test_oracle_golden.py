"""The oracle (CPU restatement) against the reference solver's own outputs.

The golden dumps were produced by the reference compiled from its sources
(tests/golden/make_golden.py).  The restatement keeps the reference's
expression and summation order, so every record must agree bit for bit."""
import os

import numpy as np
import pytest

from conftest import VARIANTS
from mel_compare import run_oracle
from mel_dump import read_dump

STEPS = [0, 1, 2, 10, 50, 200]


@pytest.mark.parametrize("name", VARIANTS)
def test_oracle_matches_reference_bitwise(name, golden_dir, tmp_path):
    model, ref = golden_dir(name)
    got, _ = run_oracle(model, STEPS, str(tmp_path / "oracle"))
    for s in STEPS:
        R = read_dump(os.path.join(ref, "step_%06d.bin" % s))
        G = got[s]
        assert set(R) <= set(G), set(R) - set(G)
        for k, v in R.items():
            w = G[k]
            assert v.shape == w.shape, (s, k)
            if v.dtype.kind == "f":
                same = np.array_equal(v.view(np.int64), w.view(np.int64)) or np.array_equal(v, w)
            else:
                same = np.array_equal(v, w)
            assert same, "step %d record %s differs" % (s, k)


def test_goldens_exercise_the_step():
    """The golden set is only a pin if contact, Neumann, Dirichlet, damping and
    cohesive-damage paths are actually non-zero somewhere in it."""
    import tempfile
    from conftest import extract_golden
    seen = {"contact": False, "neumann": False, "dirichlet": False, "damping": False, "czm": False,
            "rigid_rotation": False}
    for name in VARIANTS:
        with tempfile.TemporaryDirectory() as d:
            _, ref = extract_golden(name, d)
            for s in STEPS:
                D = read_dump(os.path.join(ref, "step_%06d.bin" % s))
                nb = int(D["n_bodies"][0])
                for b in range(nb):
                    p = "b%d." % b
                    seen["contact"] |= bool(np.any(D[p + "x_contact_force"] != 0))
                    seen["neumann"] |= bool(np.any(D[p + "y_neumann_force"] != 0))
                    seen["dirichlet"] |= bool(np.any(D[p + "x_dirichlet_force"] != 0))
                    seen["damping"] |= bool(np.any(D[p + "x_damping_force"] != 0))
                    cd = D[p + "contacts_d"].reshape(-1, 18)
                    seen["czm"] |= bool(np.any((cd[:, 17] > 0) & (cd[:, 17] < 1)))
                    seen["rigid_rotation"] |= bool(D[p + "rigid"][8] != 0)
    assert all(seen.values()), seen
